/* dns_oracle.c -- CPU FP64 restatement of the 3D channel-DNS time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never part of it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Restated from the mathematics of DNS_pressure_BC_3D.f90 (file:line cited per
 * function).  Fields are real(8) in Fortran order f(i,j,k) -> f[i + nx*(j + ny*k)];
 * spectral planes are complex (nx/2+1, ny) with kx fastest.  The 2D transform
 * follows FFTW's r2c/c2r conventions (unnormalised forward e^{-i}, backward
 * e^{+i} then divided by nx*ny, fftw3_dns_module.f90:95-132), including the
 * c2r treatment of non-Hermitian input: a full complex inverse along y for every
 * kx column, then a half-complex inverse along x that uses only the real part of
 * the kx = 0 and kx = nx/2 entries.
 *
 * Pinning (tests/test_oracle_*.py): LGL nodes/weights closed forms (n = 3, 5),
 * D1 exact on polynomials, FFT vs brute-force DFT, wavenumber and mask closed
 * forms, Poiseuille fixed point (E = 4 uc^2/15), LU residuals against the
 * explicitly assembled matrices, and the 15-step trace the reference ships in
 * standard_output.txt:88-104.
 */
#include "dns_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cpx;

static const double PI = 3.141592653589793115997963468544185161590576171875; /* 4*atan(1) */

/* ------------------------------------------------------------------------- */
/* 1D complex FFT (mixed-radix Stockham, factors 4, 2, odd primes)          */
/* ------------------------------------------------------------------------- */
typedef struct { int n, nf, f[48]; cpx* tw; } fplan;   /* tw[k] = exp(-2 pi i k/n) */

static void fplan_init(fplan* p, int n, int variant) {
    int m = n, q;
    p->n = n; p->nf = 0;
    if (variant != 1) while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
    for (q = 3; q * q <= m; q += 2) while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
    if (m > 1) p->f[p->nf++] = m;
    if (variant == 1)               /* radix 2 only, odd factors first: other rounding */
        for (int a = 0, b = p->nf - 1; a < b; a++, b--) { int t = p->f[a]; p->f[a] = p->f[b]; p->f[b] = t; }
    p->tw = (cpx*)malloc(sizeof(cpx) * (size_t)n);
    for (int k = 0; k < n; k++) {
        double ang = -2.0 * PI * (double)k / (double)n;
        p->tw[k].re = cos(ang); p->tw[k].im = sin(ang);
    }
}
static void fplan_free(fplan* p) { free(p->tw); p->tw = NULL; }

static inline cpx cmul(cpx a, cpx b) { cpx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re }; return r; }

/* sign -1: X_k = sum x_n e^{-2 pi i kn/N};  sign +1: e^{+}.  Unnormalised. */
static void fft_exec(const fplan* p, cpx* data, cpx* work, int sign) {
    int n = p->n, Ns = 1;
    cpx *a = data, *b = work;
    cpx vbuf[64], obuf[64];
    for (int s = 0; s < p->nf; s++) {
        int R = p->f[s], m = n / R;
        cpx *v = vbuf, *o = obuf;
        if (R > 64) { v = (cpx*)malloc(sizeof(cpx) * R); o = (cpx*)malloc(sizeof(cpx) * R); }
        for (int j = 0; j < m; j++) {
            int k = j % Ns;
            for (int r = 0; r < R; r++) v[r] = a[j + r * m];
            if (k) {
                int step = n / (Ns * R);
                for (int r = 1; r < R; r++) {
                    cpx w = p->tw[(k * r * step) % n];
                    if (sign > 0) w.im = -w.im;
                    v[r] = cmul(v[r], w);
                }
            }
            if (R == 2) {
                o[0].re = v[0].re + v[1].re; o[0].im = v[0].im + v[1].im;
                o[1].re = v[0].re - v[1].re; o[1].im = v[0].im - v[1].im;
            } else if (R == 4) {
                cpx s02 = { v[0].re + v[2].re, v[0].im + v[2].im };
                cpx d02 = { v[0].re - v[2].re, v[0].im - v[2].im };
                cpx s13 = { v[1].re + v[3].re, v[1].im + v[3].im };
                cpx d13 = { v[1].re - v[3].re, v[1].im - v[3].im };
                /* W = sign*i;  W*z = sign*(-z.im + i z.re) */
                cpx wd = { -sign * d13.im, sign * d13.re };
                o[0].re = s02.re + s13.re; o[0].im = s02.im + s13.im;
                o[2].re = s02.re - s13.re; o[2].im = s02.im - s13.im;
                o[1].re = d02.re + wd.re;  o[1].im = d02.im + wd.im;
                o[3].re = d02.re - wd.re;  o[3].im = d02.im - wd.im;
            } else {
                int step = n / R;
                for (int q = 0; q < R; q++) {
                    cpx acc = v[0];
                    for (int r = 1; r < R; r++) {
                        cpx w = p->tw[((q * r) % R) * step];
                        if (sign > 0) w.im = -w.im;
                        cpx t = cmul(v[r], w);
                        acc.re += t.re; acc.im += t.im;
                    }
                    o[q] = acc;
                }
            }
            int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) b[base + r * Ns] = o[r];
        }
        if (R > 64) { free(v); free(o); }
        cpx* t = a; a = b; b = t;
        Ns *= R;
    }
    if (a != data) memcpy(data, a, sizeof(cpx) * (size_t)n);
}

/* 2D plane transforms on real(nx,ny) <-> complex(nx/2+1,ny), FFTW conventions */
typedef struct { int nx, ny, nxhp; fplan px, py; } plan2d;

static void plan2d_init(plan2d* p, int nx, int ny, int variant) {
    p->nx = nx; p->ny = ny; p->nxhp = nx / 2 + 1;
    fplan_init(&p->px, nx, variant); fplan_init(&p->py, ny, variant);
}
static void plan2d_free(plan2d* p) { fplan_free(&p->px); fplan_free(&p->py); }

/* forward: fftw_execute_dft_r2c (fftw3_dns_module.f90:95-108) */
static void r2c_plane(const plan2d* p, const double* in, cpx* out, cpx* buf, cpx* work) {
    int nx = p->nx, ny = p->ny, nxhp = p->nxhp;
    for (int y = 0; y < ny; y++) {
        for (int x = 0; x < nx; x++) { buf[x].re = in[x + nx * y]; buf[x].im = 0.0; }
        fft_exec(&p->px, buf, work, -1);
        for (int k = 0; k < nxhp; k++) out[k + nxhp * y] = buf[k];
    }
    for (int k = 0; k < nxhp; k++) {
        for (int y = 0; y < ny; y++) buf[y] = out[k + nxhp * y];
        fft_exec(&p->py, buf, work, -1);
        for (int y = 0; y < ny; y++) out[k + nxhp * y] = buf[y];
    }
}

/* backward: fftw_execute_dft_c2r then divide by nx*ny (fftw3_dns_module.f90:116-132).
 * The input plane is not modified (a copy is transformed). */
static void c2r_plane(const plan2d* p, const cpx* in, double* out, cpx* tmp, cpx* buf, cpx* work) {
    int nx = p->nx, ny = p->ny, nxhp = p->nxhp;
    double norm = (double)(nx * ny);
    for (int k = 0; k < nxhp; k++) {
        for (int y = 0; y < ny; y++) buf[y] = in[k + nxhp * y];
        fft_exec(&p->py, buf, work, +1);
        for (int y = 0; y < ny; y++) tmp[k + nxhp * y] = buf[y];
    }
    for (int y = 0; y < ny; y++) {
        const cpx* g = tmp + nxhp * y;
        buf[0].re = g[0].re; buf[0].im = 0.0;
        buf[nx / 2].re = g[nx / 2].re; buf[nx / 2].im = 0.0;
        for (int k = 1; k < nx / 2; k++) {
            buf[k] = g[k];
            buf[nx - k].re = g[k].re; buf[nx - k].im = -g[k].im;
        }
        fft_exec(&p->px, buf, work, +1);
        for (int x = 0; x < nx; x++) out[x + nx * y] = buf[x].re / norm;
    }
}

/* ------------------------------------------------------------------------- */
/* LGL nodes/weights and differentiation matrix (lgl_module.f90:30-156,211-263) */
/* ------------------------------------------------------------------------- */
static void legendre_p(int n, double x, double* p) {
    if (n == 0) { *p = 1.0; return; }
    if (n == 1) { *p = x; return; }
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; k++) {
        double p2 = ((double)(2 * k - 1) * x * p1 - (double)(k - 1) * p0) / (double)k;
        p0 = p1; p1 = p2;
    }
    *p = p1;
}
static void legendre_pdd(int n, double x, double* p, double* dp, double* d2p) {
    if (n == 0) { *p = 1.0; *dp = 0.0; *d2p = 0.0; return; }
    if (n == 1) { *p = x; *dp = 1.0; *d2p = 0.0; return; }
    double p0 = 1.0, dp0 = 0.0, e0 = 0.0, p1 = x, dp1 = 1.0, e1 = 0.0;
    for (int k = 2; k <= n; k++) {
        double c = (double)(2 * k - 1), d = (double)(k - 1), kk = (double)k;
        double p2 = (c * x * p1 - d * p0) / kk;
        double dp2 = (c * (x * dp1 + p1) - d * dp0) / kk;
        double e2 = (c * (x * e1 + 2.0 * dp1) - d * e0) / kk;
        p0 = p1; p1 = p2; dp0 = dp1; dp1 = dp2; e0 = e1; e1 = e2;
    }
    *p = p1; *dp = dp1; *d2p = e1;
}

void orc_lgl(int n, double* x, double* w) {
    if (n == 1) { x[0] = 0.0; w[0] = 2.0; return; }
    if (n == 2) { x[0] = -1.0; x[1] = 1.0; w[0] = w[1] = 1.0; return; }
    x[0] = -1.0; x[n - 1] = 1.0;
    w[0] = 2.0 / (double)(n * (n - 1)); w[n - 1] = w[0];
    for (int i = 1; i < n - 1; i++) {
        double xi = -cos(PI * (double)i / (double)(n - 1));
        for (int it = 0; it < 100; it++) {
            double xo = xi, p, dp, d2p;
            legendre_pdd(n - 1, xi, &p, &dp, &d2p);
            xi = xo - dp / d2p;
            if (fabs(xi - xo) < 1.0e-15) break;
        }
        x[i] = xi;
        double p;
        legendre_p(n - 1, xi, &p);
        w[i] = 2.0 / ((double)(n * (n - 1)) * p * p);
    }
}

/* D(i,i) = sum_{k!=i} 1/(x_i-x_k); D(i,j) = prod_{k!=i,j}(x_i-x_k) / prod_{k!=j}(x_j-x_k)
 * and D(i,j) = 0 when |prod_{k!=j}(x_j-x_k)| <= 1e-15 (lgl_module.f90:23,254-258):
 * for n >= 65 that holds for every column, so the reference D1 is diagonal there. */
void orc_diff_matrix(int n, const double* x, double* d) {
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) d[i + n * j] = 0.0;
    if (n == 1) return;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            if (i == j) {
                double s = 0.0;
                for (int k = 0; k < n; k++) if (k != i) s += 1.0 / (x[i] - x[k]);
                d[i + n * i] = s;
            } else {
                double p1 = 1.0, p2 = 1.0;
                for (int k = 0; k < n; k++) {
                    if (k != i && k != j) p1 *= (x[i] - x[k]);
                    if (k != j) p2 *= (x[j] - x[k]);
                }
                d[i + n * j] = (fabs(p2) > 1.0e-15) ? p1 / p2 : 0.0;
            }
        }
    }
}

/* Exact LGL differentiation matrix (not the reference's): with N = n-1,
 * D(i,j) = P_N(x_i) / (P_N(x_j) (x_i - x_j)) for i != j, D(0,0) = -N(N+1)/4,
 * D(N,N) = N(N+1)/4, other diagonal entries 0 (the classical closed form of the Lagrange
 * derivative at the LGL nodes), evaluated in long double. */
void orc_diff_matrix_exact(int n, const double* x, double* d) {
    int N = n - 1;
    long double* p = (long double*)malloc(sizeof(long double) * n);
    for (int i = 0; i < n; i++) {
        long double p0 = 1.0L, p1 = x[i];
        if (N == 0) p1 = 1.0L;
        for (int k = 2; k <= N; k++) { long double p2 = ((2 * k - 1) * (long double)x[i] * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
        p[i] = p1;
    }
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) {
        long double v;
        if (i != j) v = p[i] / (p[j] * ((long double)x[i] - (long double)x[j]));
        else if (i == 0) v = -(long double)N * (N + 1) / 4.0L;
        else if (i == N) v = (long double)N * (N + 1) / 4.0L;
        else v = 0.0L;
        d[i + n * j] = (double)v;
    }
    free(p);
}

/* setup_spectral_3d (DNS_pressure_BC_3D.f90:769-801) */
void orc_wavenumbers(int nx, int ny, double xlen, double ylen, double* xw, double* yw) {
    double alpha = 2.0 * PI / xlen;
    for (int i = 0; i <= nx / 2; i++) xw[i] = (double)i * alpha;
    for (int j = 0; j < ny; j++) {
        if (j + 1 <= ny / 2 + 1) yw[j] = (double)j * 2.0 * PI / ylen;
        else yw[j] = (double)(j - ny) * 2.0 * PI / ylen;
    }
}

/* apply_dealias_truncation (DNS_pressure_BC_3D.f90:1163-1193): 1 = kept */
void orc_dealias_mask(int nx, int ny, unsigned char* m) {
    int nxhp = nx / 2 + 1, kxm = (nx / 2) * 2 / 3, kym = (ny / 2) * 2 / 3;
    for (int j = 0; j < ny; j++) {
        for (int i = 0; i < nxhp; i++) {
            int keep = 1;
            if (i > kxm) keep = 0;
            else if (j + 1 <= ny / 2 + 1) { if (j > kym) keep = 0; }
            else { if (abs(j - ny) > kym) keep = 0; }
            m[i + nxhp * j] = (unsigned char)keep;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Dense LU with partial pivoting (the arithmetic of zgesv on a real matrix)  */
/* ------------------------------------------------------------------------- */
static int lu_factor(int n, double* a /* col-major */, int* piv) {
    for (int k = 0; k < n; k++) {
        int p = k; double best = fabs(a[k + n * k]);
        for (int r = k + 1; r < n; r++) { double v = fabs(a[r + n * k]); if (v > best) { best = v; p = r; } }
        piv[k] = p;
        if (a[p + n * k] == 0.0) return k + 1;
        if (p != k) for (int c = 0; c < n; c++) { double t = a[k + n * c]; a[k + n * c] = a[p + n * c]; a[p + n * c] = t; }
        double rcp = 1.0 / a[k + n * k];
        for (int r = k + 1; r < n; r++) a[r + n * k] *= rcp;
        for (int c = k + 1; c < n; c++) {
            double akc = a[k + n * c];
            if (akc != 0.0) for (int r = k + 1; r < n; r++) a[r + n * c] -= a[r + n * k] * akc;
        }
    }
    return 0;
}
static void lu_solve(int n, const double* a, const int* piv, double* b) {
    for (int k = 0; k < n; k++) if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int c = 0; c < n; c++) { double bc = b[c]; if (bc != 0.0) for (int r = c + 1; r < n; r++) b[r] -= a[r + n * c] * bc; }
    for (int c = n - 1; c >= 0; c--) {
        b[c] /= a[c + n * c];
        double bc = b[c];
        if (bc != 0.0) for (int r = 0; r < c; r++) b[r] -= a[r + n * c] * bc;
    }
}

/* ------------------------------------------------------------------------- */
/* Context                                                                    */
/* ------------------------------------------------------------------------- */
struct orc_ctx {
    orc_params p;
    int nx, ny, nz, nxhp, npl, nspl;  /* npl = nx*ny, nspl = nxhp*ny */
    size_t N, NS;
    double alpha, beta, ybar, theta;
    double *xw, *yw, *xsq, *ysq, *z, *wz, *d1, *d2, *stiff;
    unsigned char* mask;
    plan2d plan;
    int nthreads;
    /* state (DNS_pressure_BC_3D.f90:113-120) */
    double *u, *v, *w, *p_total, *ustar, *vstar, *wstar, *phi, *su, *sv, *sw;
    double *gpx, *gpy;              /* grad p_total, wall planes k = 1 and nz only */
    orc_pi_state pi;
    double time; int step;
    /* scratch */
    double *f1, *f2, *f3, *f4, *f5, *f6, *f7, *f8, *f9;
    cpx *spec, *spec2;
    double *un, *vn, *wn;           /* runge_kutta_convection: velocities at the start of the step */
    long long jac_iters; int jac_fail;
    /* cached LU factors (identical to refactoring: the matrices are step-invariant) */
    double* hlu; int* hpiv; int hready;
    double* plu; int* ppiv; int pready;
    int status;
};

static double* dalloc(size_t n) { return (double*)calloc(n, sizeof(double)); }

orc_ctx* orc_create(const orc_params* p) {
    if (p->nx < 2 || p->ny < 2 || p->nz < 3 || (p->nx & 1) || (p->ny & 1)) return NULL;
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    c->p = *p;
    c->nx = p->nx; c->ny = p->ny; c->nz = p->nz; c->nxhp = p->nx / 2 + 1;
    c->npl = c->nx * c->ny; c->nspl = c->nxhp * c->ny;
    c->N = (size_t)c->npl * c->nz; c->NS = (size_t)c->nspl * c->nz;
    /* setup_grid_3d (DNS_pressure_BC_3D.f90:740-742): alpha, beta recomputed, ybar forced */
    c->alpha = 2.0 * PI / p->xlen; c->beta = 2.0 * PI / p->ylen; c->ybar = 2.0;
    c->theta = p->use_crank_nicolson ? 0.5 : 1.0;
    int nx = c->nx, ny = c->ny, nz = c->nz, nxhp = c->nxhp;
    c->xw = dalloc(nxhp); c->xsq = dalloc(nxhp); c->yw = dalloc(ny); c->ysq = dalloc(ny);
    orc_wavenumbers(nx, ny, p->xlen, p->ylen, c->xw, c->yw);
    for (int i = 0; i < nxhp; i++) c->xsq[i] = c->xw[i] * c->xw[i];
    for (int j = 0; j < ny; j++) c->ysq[j] = c->yw[j] * c->yw[j];
    c->z = dalloc(nz); c->wz = dalloc(nz); c->d1 = dalloc((size_t)nz * nz); c->d2 = dalloc((size_t)nz * nz);
    orc_lgl(nz, c->z, c->wz);
    if (p->lgl_derivative) orc_diff_matrix_exact(nz, c->z, c->d1);
    else orc_diff_matrix(nz, c->z, c->d1);
    /* d2 = matmul(d1, d1) (DNS_pressure_BC_3D.f90:801) */
    for (int j = 0; j < nz; j++) for (int i = 0; i < nz; i++) {
        double s = 0.0;
        for (int k = 0; k < nz; k++) s += c->d1[i + nz * k] * c->d1[k + nz * j];
        c->d2[i + nz * j] = s;
    }
    /* stiffness(k,n) = sum_m d1(m,k) d1(m,n) w_m * (2/ybar) (DNS_pressure_BC_3D.f90:1464-1471) */
    c->stiff = dalloc((size_t)nz * nz);
    for (int n = 0; n < nz; n++) for (int k = 0; k < nz; k++) {
        double s = 0.0;
        for (int m = 0; m < nz; m++) s += c->d1[m + nz * k] * c->d1[m + nz * n] * c->wz[m];
        c->stiff[k + nz * n] = s * (2.0 / c->ybar);
    }
    c->mask = (unsigned char*)malloc((size_t)c->nspl);
    orc_dealias_mask(nx, ny, c->mask);
    plan2d_init(&c->plan, nx, ny, p->fft_variant);
#ifdef _OPENMP
    c->nthreads = omp_get_max_threads();
#else
    c->nthreads = 1;
#endif
    size_t N = c->N;
    c->u = dalloc(N); c->v = dalloc(N); c->w = dalloc(N); c->p_total = dalloc(N);
    c->ustar = dalloc(N); c->vstar = dalloc(N); c->wstar = dalloc(N); c->phi = dalloc(N);
    c->su = dalloc(N); c->sv = dalloc(N); c->sw = dalloc(N);
    c->gpx = dalloc((size_t)2 * c->npl); c->gpy = dalloc((size_t)2 * c->npl);
    c->f1 = dalloc(N); c->f2 = dalloc(N); c->f3 = dalloc(N); c->f4 = dalloc(N); c->f5 = dalloc(N);
    c->f6 = dalloc(N); c->f7 = dalloc(N); c->f8 = dalloc(N); c->f9 = dalloc(N);
    c->spec = (cpx*)calloc(c->NS, sizeof(cpx)); c->spec2 = (cpx*)calloc(c->NS, sizeof(cpx));
    if (p->convection_scheme == 1) { c->un = dalloc(N); c->vn = dalloc(N); c->wn = dalloc(N); }
    /* read_input_3d (DNS_pressure_BC_3D.f90:629-641) */
    c->pi.G = p->target_pressure_gradient; c->pi.gavg = p->target_pressure_gradient;
    return c;
}

void orc_destroy(orc_ctx* c) {
    if (!c) return;
    double* fs[] = { c->xw, c->xsq, c->yw, c->ysq, c->z, c->wz, c->d1, c->d2, c->stiff, c->u, c->v, c->w,
                     c->p_total, c->ustar, c->vstar, c->wstar, c->phi, c->su, c->sv, c->sw, c->gpx, c->gpy,
                     c->f1, c->f2, c->f3, c->f4, c->f5, c->f6, c->f7, c->f8, c->f9, c->hlu, c->plu,
                     c->un, c->vn, c->wn };
    for (size_t i = 0; i < sizeof(fs) / sizeof(fs[0]); i++) free(fs[i]);
    free(c->mask); free(c->spec); free(c->spec2); free(c->hpiv); free(c->ppiv);
    plan2d_free(&c->plan);
    free(c);
}

/* per-thread FFT scratch */
typedef struct { cpx *tmp, *buf, *work; } fscratch;
static void fs_alloc(orc_ctx* c, fscratch* s) {
    int m = c->nx > c->ny ? c->nx : c->ny;
    s->tmp = (cpx*)malloc(sizeof(cpx) * (size_t)c->nspl);
    s->buf = (cpx*)malloc(sizeof(cpx) * (size_t)m);
    s->work = (cpx*)malloc(sizeof(cpx) * (size_t)m);
}
static void fs_free(fscratch* s) { free(s->tmp); free(s->buf); free(s->work); }

/* forward transform of every z-plane */
static void r2c_field(orc_ctx* c, const double* f, cpx* out) {
    #pragma omp parallel
    {
        fscratch s; fs_alloc(c, &s);
        #pragma omp for schedule(static)
        for (int k = 0; k < c->nz; k++)
            r2c_plane(&c->plan, f + (size_t)c->npl * k, out + (size_t)c->nspl * k, s.buf, s.work);
        fs_free(&s);
    }
}
static void c2r_field(orc_ctx* c, const cpx* in, double* f) {
    #pragma omp parallel
    {
        fscratch s; fs_alloc(c, &s);
        #pragma omp for schedule(static)
        for (int k = 0; k < c->nz; k++)
            c2r_plane(&c->plan, in + (size_t)c->nspl * k, f + (size_t)c->npl * k, s.tmp, s.buf, s.work);
        fs_free(&s);
    }
}

/* compute_derivatives_3d (DNS_pressure_BC_3D.f90:2321-2407): any of dx, dy, dz (NULL = skip) */
static void derivatives(orc_ctx* c, const double* f, double* dx, double* dy, double* dz) {
    int nx = c->nx, ny = c->ny, nz = c->nz, nxhp = c->nxhp, npl = c->npl, nspl = c->nspl;
    if (dz) {
        #pragma omp parallel for schedule(static)
        for (int j = 0; j < ny; j++) {
            for (int i = 0; i < nx; i++) {
                size_t col = (size_t)i + (size_t)nx * j;
                for (int k = 0; k < nz; k++) {
                    double s = 0.0;
                    for (int m = 0; m < nz; m++) s += c->d1[k + nz * m] * f[col + (size_t)npl * m];
                    dz[col + (size_t)npl * k] = (2.0 / c->ybar) * s;
                }
            }
        }
    }
    if (dx || dy) {
        #pragma omp parallel
        {
            fscratch s; fs_alloc(c, &s);
            cpx* fh = (cpx*)malloc(sizeof(cpx) * (size_t)nspl);
            cpx* dh = (cpx*)malloc(sizeof(cpx) * (size_t)nspl);
            #pragma omp for schedule(static)
            for (int k = 0; k < nz; k++) {
                r2c_plane(&c->plan, f + (size_t)npl * k, fh, s.buf, s.work);
                if (dx) {
                    for (int j = 0; j < ny; j++) for (int i = 0; i < nxhp; i++) {
                        cpx a = fh[i + nxhp * j];
                        dh[i + nxhp * j].re = a.re * 0.0 - a.im * c->xw[i];
                        dh[i + nxhp * j].im = a.re * c->xw[i] + a.im * 0.0;
                    }
                    c2r_plane(&c->plan, dh, dx + (size_t)npl * k, s.tmp, s.buf, s.work);
                }
                if (dy) {
                    for (int j = 0; j < ny; j++) for (int i = 0; i < nxhp; i++) {
                        cpx a = fh[i + nxhp * j];
                        dh[i + nxhp * j].re = a.re * 0.0 - a.im * c->yw[j];
                        dh[i + nxhp * j].im = a.re * c->yw[j] + a.im * 0.0;
                    }
                    c2r_plane(&c->plan, dh, dy + (size_t)npl * k, s.tmp, s.buf, s.work);
                }
            }
            free(fh); free(dh); fs_free(&s);
        }
    }
}

/* calculate_bulk_velocity_3d (DNS_pressure_BC_3D.f90:2434-2464) */
static double bulk_of(int nx, int ny, int nz, const double* u, const double* wz) {
    double vol = 0.0, ws = 0.0;
    for (int k = 0; k < nz; k++) {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                vol += u[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] * wz[k];
        ws += wz[k];
    }
    return vol / ((double)(nx * ny) * ws);
}
static double bulk_velocity(const orc_ctx* c) { return bulk_of(c->nx, c->ny, c->nz, c->u, c->wz); }

double orc_bulk_velocity(int nx, int ny, int nz, const double* u) {
    double* z = dalloc(nz); double* wz = dalloc(nz);
    orc_lgl(nz, z, wz);
    double ub = bulk_of(nx, ny, nz, u, wz);
    free(z); free(wz);
    return ub;
}

/* update_pressure_gradient_3d (DNS_pressure_BC_3D.f90:2488-2549) */
int orc_pi_update(const orc_params* p, orc_pi_state* s, int istep, double ub) {
    if (istep % p->controller_update_freq != 0) return 0;
    s->Ub = ub;
    s->err = p->target_bulk_velocity - s->Ub;
    double kp = p->controller_gain, ki = 0.05 * p->controller_gain;
    double base = 3.0 / p->re;
    double prop = kp * s->err, integ = ki * s->err_int;
    double gu = base + prop + integ;
    double gc = fmax(0.0, fmin(10.0, gu));
    if (fabs(gc - gu) < 1.0e-12) {                       /* anti-windup: frozen while clamped */
        if (istep <= p->controller_update_freq) s->err_int = s->err_int + s->err * p->dt;   /* Euler */
        else s->err_int = s->err_int + 0.5 * p->dt * (s->err_prev + s->err);               /* trapezoid */
    }
    s->err_prev = s->err;
    s->G = gc;
    s->gsum += s->G; s->gcount += 1;
    s->gavg = s->gsum / (double)s->gcount;
    return 1;
}

/* called with istart, not the step number (DNS_pressure_BC_3D.f90:1118) */
static void update_pressure_gradient(orc_ctx* c, int istep) {
    if (istep % c->p.controller_update_freq != 0) return;
    orc_pi_update(&c->p, &c->pi, istep, bulk_velocity(c));
}

static void dealias_field(orc_ctx* c, double* f) {
    int nspl = c->nspl, npl = c->npl;
    #pragma omp parallel
    {
        fscratch s; fs_alloc(c, &s);
        cpx* fh = (cpx*)malloc(sizeof(cpx) * (size_t)nspl);
        #pragma omp for schedule(static)
        for (int k = 0; k < c->nz; k++) {
            r2c_plane(&c->plan, f + (size_t)npl * k, fh, s.buf, s.work);
            for (int m = 0; m < nspl; m++) if (!c->mask[m]) { fh[m].re = 0.0; fh[m].im = 0.0; }
            c2r_plane(&c->plan, fh, f + (size_t)npl * k, s.tmp, s.buf, s.work);
        }
        free(fh); fs_free(&s);
    }
}

/* su, sv, sw = u x omega of c->u, c->v, c->w, dealiased when on
 * (compute_source_terms_3d_standard :1054-1098 / _dealiased :1233-1261) */
static void rotational_product(orc_ctx* c) {
    size_t N = c->N;
    double *ux = c->f1, *uy = c->f2, *uz = c->f3, *vx = c->f4, *vy = c->f5, *vz = c->f6;
    double *wx = c->f7, *wy = c->f8, *wz = c->f9;
    derivatives(c, c->u, ux, uy, uz);
    derivatives(c, c->v, vx, vy, vz);
    derivatives(c, c->w, wx, wy, wz);
    /* omega (:1077-1079) */
    #pragma omp parallel for schedule(static)
    for (size_t n = 0; n < N; n++) {
        double ox = wy[n] - vz[n], oy = uz[n] - wx[n], oz = vx[n] - uy[n];
        double u = c->u[n], v = c->v[n], w = c->w[n];
        c->su[n] = v * oz - w * oy;
        c->sv[n] = w * ox - u * oz;
        c->sw[n] = u * oy - v * ox;
    }
    if (c->p.use_dealias_23_rule) {
        dealias_field(c, c->su); dealias_field(c, c->sv); dealias_field(c, c->sw);
    }
}

/* compute_source_terms_3d (DNS_pressure_BC_3D.f90:1032-1148).  With convection_scheme 1 the
 * product is staged as runge_kutta_convection does (DNS_pressure_BC_2D.f90:931-1014): from
 * un = u, for irk = 4, 3, 2, 1: S = S(u), u = un + (1/irk) dt S, walls zero; the S of the last
 * stage is the source term and u is reset to un.  The forcing follows, as in the 2D loop
 * (apply_flow_control after convection_step, DNS_pressure_BC_2D.f90:217-223). */
void orc_compute_source_terms(orc_ctx* c) {
    size_t N = c->N;
    if (c->p.convection_scheme == 1) {
        size_t b = sizeof(double) * N, top = (size_t)c->npl * (c->nz - 1);
        memcpy(c->un, c->u, b); memcpy(c->vn, c->v, b); memcpy(c->wn, c->w, b);
        for (int irk = 4; irk >= 1; irk--) {
            double alf = 1.0 / (double)irk, dt = c->p.dt;
            rotational_product(c);
            for (size_t n = 0; n < N; n++) {
                c->u[n] = c->un[n] + alf * dt * c->su[n];
                c->v[n] = c->vn[n] + alf * dt * c->sv[n];
                c->w[n] = c->wn[n] + alf * dt * c->sw[n];
            }
            for (int m = 0; m < c->npl; m++) {      /* apply_wall_bc after each stage */
                c->u[m] = 0.0; c->v[m] = 0.0; c->w[m] = 0.0;
                c->u[top + m] = 0.0; c->v[top + m] = 0.0; c->w[top + m] = 0.0;
            }
        }
        memcpy(c->u, c->un, b); memcpy(c->v, c->vn, b); memcpy(c->w, c->wn, b);
    } else {
        rotational_product(c);
    }
    /* apply_flow_control_forcing (:1107-1148) */
    if (c->p.flow_control_method == 1) c->pi.G = c->p.target_pressure_gradient;
    else update_pressure_gradient(c, c->p.istart);
    double G = c->pi.G;
    for (size_t n = 0; n < N; n++) c->su[n] = c->su[n] + G;
}

/* compute_laplacian_3d (DNS_pressure_BC_3D.f90:1564-1606) into out */
static void laplacian(orc_ctx* c, const double* f, double* out) {
    int nz = c->nz, nxhp = c->nxhp, ny = c->ny, nspl = c->nspl;
    double zs = 2.0 / c->ybar;
    r2c_field(c, f, c->spec);
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++) {
        double vr[512], vi[512];
        for (int i = 0; i < nxhp; i++) {
            size_t m0 = (size_t)i + (size_t)nxhp * j;
            double ks = c->xsq[i] + c->ysq[j];
            for (int k = 0; k < nz; k++) {
                double sr = 0.0, si = 0.0;
                for (int m = 0; m < nz; m++) {
                    cpx a = c->spec[m0 + (size_t)nspl * m];
                    sr += c->d2[k + nz * m] * a.re; si += c->d2[k + nz * m] * a.im;
                }
                vr[k] = sr * (zs * zs); vi[k] = si * (zs * zs);
            }
            for (int k = 0; k < nz; k++) {
                cpx a = c->spec[m0 + (size_t)nspl * k];
                c->spec2[m0 + (size_t)nspl * k].re = vr[k] - ks * a.re;
                c->spec2[m0 + (size_t)nspl * k].im = vi[k] - ks * a.im;
            }
        }
    }
    c2r_field(c, c->spec2, out);
}

static void helmholtz_matrix(const orc_ctx* c, double factor, int i, int j, double* h) {
    int nz = c->nz;
    for (int n = 0; n < nz * nz; n++) h[n] = c->stiff[n] * factor;
    for (int k = 0; k < nz; k++)
        h[k + nz * k] = h[k + nz * k] + (1.0 + factor * (c->xsq[i] + c->ysq[j])) * c->wz[k] * (c->ybar / 2.0);
}

/* solve_viscous_helmholtz (DNS_pressure_BC_3D.f90:1425-1521) */
static void helmholtz(orc_ctx* c, double* fout, const double* rhs, double theta, const double* bc1, const double* bc2) {
    int nz = c->nz, nxhp = c->nxhp, nspl = c->nspl;
    double factor = theta * c->p.dt / c->p.re;
    r2c_field(c, rhs, c->spec);
    cpx* bh1 = (cpx*)malloc(sizeof(cpx) * (size_t)nspl);
    cpx* bh2 = (cpx*)malloc(sizeof(cpx) * (size_t)nspl);
    {
        fscratch s; fs_alloc(c, &s);
        r2c_plane(&c->plan, bc1, bh1, s.buf, s.work);
        r2c_plane(&c->plan, bc2, bh2, s.buf, s.work);
        fs_free(&s);
    }
    int cached = !c->p.refactor_every_step;
    if (cached && !c->hready) {
        c->hlu = dalloc((size_t)nz * nz * nspl); c->hpiv = (int*)calloc((size_t)nz * nspl, sizeof(int));
    }
    /* the reference's mode loop is serial (:1475); with cached factors the modes are
     * independent, so they run in parallel (bitwise the same per mode) */
    const int was_ready = c->hready;
    #pragma omp parallel if (cached)
    {
    double* hm = dalloc((size_t)nz * nz);
    int* piv = (int*)calloc(nz, sizeof(int));
    double* br = dalloc(nz); double* bi = dalloc(nz);
    #pragma omp for schedule(dynamic, 8)
    for (int mi = 0; mi < nspl; mi++) {
        const int i = mi % nxhp, j = mi / nxhp;
        {
            size_t m0 = (size_t)i + (size_t)nxhp * j;
            double* a = cached ? c->hlu + (size_t)nz * nz * m0 : hm;
            int* pv = cached ? c->hpiv + (size_t)nz * m0 : piv;
            helmholtz_matrix(c, factor, i, j, hm);
            for (int k = 0; k < nz; k++) {
                cpx r = c->spec[m0 + (size_t)nspl * k];
                br[k] = r.re * c->wz[k] * (c->ybar / 2.0);
                bi[k] = r.im * c->wz[k] * (c->ybar / 2.0);
            }
            /* lifting with the assembled matrix; rows 1, nz become unit rows but columns 1, nz
             * stay, so the boundary value is effectively subtracted twice (:1494-1502) */
            for (int k = 1; k < nz - 1; k++) {
                double h1 = hm[k], hn = hm[k + nz * (nz - 1)];
                br[k] = br[k] - (h1 * bh1[m0].re) - (hn * bh2[m0].re);
                bi[k] = bi[k] - (h1 * bh1[m0].im) - (hn * bh2[m0].im);
            }
            br[0] = bh1[m0].re; bi[0] = bh1[m0].im;
            br[nz - 1] = bh2[m0].re; bi[nz - 1] = bh2[m0].im;
            if (!cached || !was_ready) {
                for (int n = 0; n < nz; n++) { hm[0 + nz * n] = 0.0; hm[(nz - 1) + nz * n] = 0.0; }
                hm[0] = 1.0; hm[(nz - 1) + nz * (nz - 1)] = 1.0;
                if (cached) memcpy(a, hm, sizeof(double) * (size_t)nz * nz);
                int info = lu_factor(nz, a, pv);
                if (info) {
                    #pragma omp critical
                    { if (!c->status) c->status = info; }   /* the reference stops here */
                }
            }
            lu_solve(nz, a, pv, br); lu_solve(nz, a, pv, bi);
            for (int k = 0; k < nz; k++) { c->spec2[m0 + (size_t)nspl * k].re = br[k]; c->spec2[m0 + (size_t)nspl * k].im = bi[k]; }
        }
    }
    free(hm); free(piv); free(br); free(bi);
    }
    if (cached) c->hready = 1;
    c2r_field(c, c->spec2, fout);
    free(bh1); free(bh2);
}

/* viscous_step_3d (DNS_pressure_BC_3D.f90:1285-1332) */
void orc_viscous_step(orc_ctx* c) {
    size_t N = c->N; int npl = c->npl, nz = c->nz;
    double dt = c->p.dt, theta = c->theta;
    double *ru = c->f1, *rv = c->f2, *rw = c->f3, *lap = c->f4;
    for (size_t n = 0; n < N; n++) { ru[n] = c->u[n] + dt * c->su[n]; rv[n] = c->v[n] + dt * c->sv[n]; rw[n] = c->w[n] + dt * c->sw[n]; }
    if (theta < 1.0) {
        double factor = 0.5 * dt / c->p.re;
        laplacian(c, c->u, lap); for (size_t n = 0; n < N; n++) ru[n] = ru[n] + factor * lap[n];
        laplacian(c, c->v, lap); for (size_t n = 0; n < N; n++) rv[n] = rv[n] + factor * lap[n];
        laplacian(c, c->w, lap); for (size_t n = 0; n < N; n++) rw[n] = rw[n] + factor * lap[n];
    }
    double* b1 = dalloc(npl); double* b2 = dalloc(npl); double* zero = dalloc(npl);
    for (int m = 0; m < npl; m++) { b1[m] = -dt * c->gpx[m]; b2[m] = -dt * c->gpx[npl + m]; }
    helmholtz(c, c->ustar, ru, theta, b1, b2);
    for (int m = 0; m < npl; m++) { b1[m] = -dt * c->gpy[m]; b2[m] = -dt * c->gpy[npl + m]; }
    helmholtz(c, c->vstar, rv, theta, b1, b2);
    helmholtz(c, c->wstar, rw, theta, zero, zero);
    free(b1); free(b2); free(zero);
    (void)nz;
}

/* compute_divergence_3d (DNS_pressure_BC_3D.f90:1858-1879) */
static void divergence(orc_ctx* c, const double* fx, const double* fy, const double* fz, double* div) {
    double *dyf = c->f8, *dzf = c->f9;
    derivatives(c, fx, div, NULL, NULL);
    derivatives(c, fy, NULL, dyf, NULL);
    derivatives(c, fz, NULL, NULL, dzf);
    for (size_t n = 0; n < c->N; n++) div[n] = div[n] + dyf[n] + dzf[n];
}

/* jacgs (DNS_pressure_BC_2D.f90:2361-2476): Jacobi-preconditioned conjugate gradients with the
 * reference's loop order, small = 1e-30 guards and absolute stop ||r||_2 <= tol (checked on the
 * initial residual and after each iteration). */
int orc_jacgs(int n, const double* a, double* x, const double* b, int niter, double tol, int* iters) {
    const double small = 1.0e-30;
    double* r = dalloc(n); double* p = dalloc(n); double* q = dalloc(n);
    double* apn = dalloc(n); double* diag = dalloc(n); double* sv = dalloc(n);
    int info = 0, it = 0;
    for (int i = 0; i < n; i++) {
        diag[i] = fabs(a[i + (size_t)n * i]) > small ? 1.0 / a[i + (size_t)n * i] : 1.0;
        sv[i] = 0.0;
    }
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) sv[i] = sv[i] + a[i + (size_t)n * j] * x[j];
    for (int i = 0; i < n; i++) { r[i] = b[i] - sv[i]; q[i] = r[i] * diag[i]; p[i] = q[i]; }
    double res0 = 0.0;
    for (int i = 0; i < n; i++) res0 = res0 + r[i] * r[i];
    res0 = sqrt(res0);
    if (!(res0 <= tol)) {
        double res1 = res0;
        info = -1;
        for (it = 1; it <= niter; it++) {
            for (int i = 0; i < n; i++) apn[i] = 0.0;
            for (int j = 0; j < n; j++) for (int i = 0; i < n; i++) apn[i] = apn[i] + a[i + (size_t)n * j] * p[j];
            double qdr = 0.0, pdapn = 0.0;
            for (int i = 0; i < n; i++) { qdr = qdr + q[i] * r[i]; pdapn = pdapn + p[i] * apn[i]; }
            double alfa = qdr / (pdapn + small);
            for (int i = 0; i < n; i++) { x[i] = x[i] + alfa * p[i]; r[i] = r[i] - alfa * apn[i]; q[i] = r[i] * diag[i]; }
            double qdrnew = 0.0;
            for (int i = 0; i < n; i++) qdrnew = qdrnew + q[i] * r[i];
            double beta = qdrnew / (qdr + small);
            for (int i = 0; i < n; i++) p[i] = r[i] * diag[i] + beta * p[i];
            res1 = 0.0;
            for (int i = 0; i < n; i++) res1 = res1 + r[i] * r[i];
            res1 = sqrt(res1);
            if (res1 <= tol) { info = 0; break; }
        }
        if (it > niter) it = niter;
    }
    if (iters) *iters = it;
    free(r); free(p); free(q); free(apn); free(diag); free(sv);
    return info;
}

void orc_jacgs_stats(const orc_ctx* c, long long* iterations, int* nonconverged) {
    if (iterations) *iterations = c->jac_iters;
    if (nonconverged) *nonconverged = c->jac_fail;
}

/* solve_pressure_3d_standard (DNS_pressure_BC_3D.f90:1654-1733), or with
 * use_weak_form_divergence the RHS of solve_pressure_3d_weak_form (:1749-1835): per mode
 * rhs_k = [-(i xw u^ + i yw v^)_k w_k ybar/2 - sum_m D1(m,k) w^_m w_m 2/ybar] / dt from the
 * transformed u*, v*, w* (:1789-1802), with the same matrix and mode-(1,1) treatment. */
void orc_solve_pressure(orc_ctx* c) {
    int nz = c->nz, nxhp = c->nxhp, nspl = c->nspl;
    const int weak = c->p.use_weak_form_divergence;
    double* div = c->f7;
    cpx *vh = NULL, *wh = NULL;
    if (weak) {
        vh = (cpx*)calloc(c->NS, sizeof(cpx)); wh = (cpx*)calloc(c->NS, sizeof(cpx));
        r2c_field(c, c->ustar, c->spec);
        r2c_field(c, c->vstar, vh);
        r2c_field(c, c->wstar, wh);
    } else {
        divergence(c, c->ustar, c->vstar, c->wstar, div);
        for (size_t n = 0; n < c->N; n++) div[n] = div[n] / c->p.dt;
        r2c_field(c, div, c->spec);
    }
    int cached = !c->p.refactor_every_step;
    const int jac = c->p.pressure_solver == 1;
    c->jac_iters = 0; c->jac_fail = 0;
    if (cached && !c->pready) {
        c->plu = dalloc((size_t)nz * nz * nspl); c->ppiv = (int*)calloc((size_t)nz * nspl, sizeof(int));
    }
    /* serial in the reference (:1680); parallel over independent modes with cached factors */
    const int was_ready = c->pready;
    #pragma omp parallel if (cached)
    {
    double* pm = dalloc((size_t)nz * nz);
    int* piv = (int*)calloc(nz, sizeof(int));
    double* br = dalloc(nz); double* bi = dalloc(nz);
    #pragma omp for schedule(dynamic, 8)
    for (int mi = 0; mi < nspl; mi++) {
        const int i = mi % nxhp, j = mi / nxhp;
        {
            size_t m0 = (size_t)i + (size_t)nxhp * j;
            double* a = cached ? c->plu + (size_t)nz * nz * m0 : pm;
            int* pv = cached ? c->ppiv + (size_t)nz * m0 : piv;
            if (!cached || !was_ready) {
                /* the O(nz^3) rebuild per mode (:1684-1692) */
                for (int k = 0; k < nz; k++) {
                    for (int n = 0; n < nz; n++) {
                        double s = 0.0;
                        for (int m = 0; m < nz; m++) s += c->d1[m + nz * k] * c->d1[m + nz * n] * c->wz[m];
                        pm[k + nz * n] = -s * (2.0 / c->ybar);
                    }
                    pm[k + nz * k] = pm[k + nz * k] - (c->xsq[i] + c->ysq[j]) * c->wz[k] * (c->ybar / 2.0);
                }
                if (i == 0 && j == 0) {
                    for (int k = 0; k < nz; k++) { pm[0 + nz * k] = 0.0; pm[k] = 0.0; }
                    pm[0] = 1.0;
                }
                if (cached) memcpy(a, pm, sizeof(double) * (size_t)nz * nz);
                if (!jac) {
                    int info = lu_factor(nz, a, pv);
                    if (info) {
                        #pragma omp critical
                        { if (!c->status) c->status = info; }   /* the reference stops here */
                    }
                }
            }
            if (weak) {
                const double kx = c->xw[i], ky = c->yw[j];
                for (int k = 0; k < nz; k++) {
                    const cpx u = c->spec[m0 + (size_t)nspl * k], v = vh[m0 + (size_t)nspl * k];
                    /* cmplx(0, xw) u^ + cmplx(0, yw) v^, negated, then times w_k and ybar/2 */
                    const double hr = (0.0 * u.re - kx * u.im) + (0.0 * v.re - ky * v.im);
                    const double hi = (0.0 * u.im + kx * u.re) + (0.0 * v.im + ky * v.re);
                    double rr = (-hr) * c->wz[k] * (c->ybar / 2.0), ri = (-hi) * c->wz[k] * (c->ybar / 2.0);
                    double sr = 0.0, si = 0.0;
                    for (int m = 0; m < nz; m++) {
                        const cpx wm = wh[m0 + (size_t)nspl * m];
                        sr += c->d1[m + nz * k] * wm.re * c->wz[m];
                        si += c->d1[m + nz * k] * wm.im * c->wz[m];
                    }
                    rr = rr - sr * (2.0 / c->ybar);
                    ri = ri - si * (2.0 / c->ybar);
                    br[k] = rr / c->p.dt; bi[k] = ri / c->p.dt;
                }
            } else {
                for (int k = 0; k < nz; k++) {
                    cpx r = c->spec[m0 + (size_t)nspl * k];
                    double sc = 0.5 * c->ybar * c->wz[k];
                    br[k] = r.re * sc; bi[k] = r.im * sc;
                }
            }
            if (i == 0 && j == 0) { br[0] = 0.0; bi[0] = 0.0; }
            if (jac) {
                /* jacgs on the real and the imaginary part, each from x = 0 (DNS_pressure_BC_2D.f90:1629) */
                double* xr = dalloc(nz); double* xi = dalloc(nz);
                int ir = 0, ii = 0;
                int fr = orc_jacgs(nz, a, xr, br, 500, c->p.cgstol, &ir);
                int fi = orc_jacgs(nz, a, xi, bi, 500, c->p.cgstol, &ii);
                #pragma omp atomic
                c->jac_iters += ir + ii;
                if (fr || fi) {
                    #pragma omp atomic
                    c->jac_fail += (fr != 0) + (fi != 0);
                }
                memcpy(br, xr, sizeof(double) * nz); memcpy(bi, xi, sizeof(double) * nz);
                free(xr); free(xi);
            } else {
                lu_solve(nz, a, pv, br); lu_solve(nz, a, pv, bi);
            }
            for (int k = 0; k < nz; k++) { c->spec2[m0 + (size_t)nspl * k].re = br[k]; c->spec2[m0 + (size_t)nspl * k].im = bi[k]; }
        }
    }
    free(pm); free(piv); free(br); free(bi);
    }
    if (cached) c->pready = 1;
    c2r_field(c, c->spec2, c->phi);
    free(vh); free(wh);
}

/* pressure_correction_3d (DNS_pressure_BC_3D.f90:1903-1929); grad p_total is kept
 * on the two wall planes only, the only ones read (:1325-1326) */
void orc_pressure_correction(orc_ctx* c) {
    size_t N = c->N; int npl = c->npl, nz = c->nz;
    double dt = c->p.dt;
    double *px = c->f1, *py = c->f2, *pz = c->f3;
    derivatives(c, c->phi, px, py, pz);
    for (size_t n = 0; n < N; n++) {
        c->u[n] = c->ustar[n] - dt * px[n];
        c->v[n] = c->vstar[n] - dt * py[n];
        c->w[n] = c->wstar[n] - dt * pz[n];
        c->p_total[n] = c->p_total[n] + c->phi[n];
    }
    fscratch s; fs_alloc(c, &s);
    cpx* fh = (cpx*)malloc(sizeof(cpx) * (size_t)c->nspl);
    cpx* dh = (cpx*)malloc(sizeof(cpx) * (size_t)c->nspl);
    for (int wall = 0; wall < 2; wall++) {
        int k = wall ? nz - 1 : 0;
        r2c_plane(&c->plan, c->p_total + (size_t)npl * k, fh, s.buf, s.work);
        for (int dir = 0; dir < 2; dir++) {
            for (int j = 0; j < c->ny; j++) for (int i = 0; i < c->nxhp; i++) {
                cpx a = fh[i + c->nxhp * j];
                double kk = dir ? c->yw[j] : c->xw[i];
                dh[i + c->nxhp * j].re = a.re * 0.0 - a.im * kk;
                dh[i + c->nxhp * j].im = a.re * kk + a.im * 0.0;
            }
            c2r_plane(&c->plan, dh, (dir ? c->gpy : c->gpx) + (size_t)npl * wall, s.tmp, s.buf, s.work);
        }
    }
    free(fh); free(dh); fs_free(&s);
}

/* apply_boundary_conditions_3d (DNS_pressure_BC_3D.f90:1953-1972) */
void orc_apply_boundary_conditions(orc_ctx* c) {
    size_t top = (size_t)c->npl * (c->nz - 1);
    for (int m = 0; m < c->npl; m++) {
        c->u[m] = 0.0; c->v[m] = 0.0; c->w[m] = 0.0;
        c->u[top + m] = 0.0; c->v[top + m] = 0.0; c->w[top + m] = 0.0;
    }
}

/* the time loop body (DNS_pressure_BC_3D.f90:286-326) */
void orc_step(orc_ctx* c, int nsteps) {
    for (int s = 0; s < nsteps; s++) {
        c->step += 1;
        c->time = c->time + c->p.dt;
        orc_compute_source_terms(c);
        orc_viscous_step(c);
        orc_solve_pressure(c);
        orc_pressure_correction(c);
        orc_apply_boundary_conditions(c);
    }
}

/* output_diagnostics_3d (DNS_pressure_BC_3D.f90:2094-2146) */
void orc_diagnostics(orc_ctx* c, orc_diag* d) {
    size_t N = c->N;
    double um = 0, vm = 0, wm = 0;
    for (size_t n = 0; n < N; n++) {
        if (fabs(c->u[n]) > um) um = fabs(c->u[n]);
        if (fabs(c->v[n]) > vm) vm = fabs(c->v[n]);
        if (fabs(c->w[n]) > wm) wm = fabs(c->w[n]);
    }
    double e = 0.0;
    double f = (c->p.xlen / (double)c->nx) * (c->p.ylen / (double)c->ny) * (c->ybar / 2.0);
    for (int i = 0; i < c->nx; i++) for (int j = 0; j < c->ny; j++) for (int k = 0; k < c->nz; k++) {
        size_t n = (size_t)i + (size_t)c->nx * ((size_t)j + (size_t)c->ny * k);
        double u = c->u[n], v = c->v[n], w = c->w[n];
        e = e + 0.5 * (u * u + v * v + w * w) * f * c->wz[k];
    }
    e = e / (c->p.xlen * c->p.ylen * c->ybar);
    d->umax = um; d->vmax = vm; d->wmax = wm; d->energy = e;
    d->pressure_gradient = c->pi.G;
    d->bulk_velocity = bulk_velocity(c);
    if (c->p.flow_control_method == 2) c->pi.Ub = d->bulk_velocity;
    d->avg_pressure_gradient = c->pi.gavg;
    d->bulk_error_integral = c->pi.err_int; d->bulk_error_prev = c->pi.err_prev;
    d->time = c->time; d->step = c->step;
}

/* check_divergence_3d (DNS_pressure_BC_3D.f90:2171-2190) */
double orc_max_divergence(orc_ctx* c) {
    double* div = c->f7;
    divergence(c, c->u, c->v, c->w, div);
    double m = 0.0;
    for (size_t n = 0; n < c->N; n++) if (fabs(div[n]) > m) m = fabs(div[n]);
    return m;
}

const double* orc_source_terms(const orc_ctx* c, int comp) { return comp == 0 ? c->su : comp == 1 ? c->sv : c->sw; }
const double* orc_star_fields(const orc_ctx* c, int comp) { return comp == 0 ? c->ustar : comp == 1 ? c->vstar : c->wstar; }
const double* orc_phi(const orc_ctx* c) { return c->phi; }
int orc_status(const orc_ctx* c) { return c->status; }

int orc_set_fields(orc_ctx* c, const double* u, const double* v, const double* w, const double* p) {
    size_t b = sizeof(double) * c->N;
    if (u) memcpy(c->u, u, b);
    if (v) memcpy(c->v, v, b);
    if (w) memcpy(c->w, w, b);
    if (p) memcpy(c->p_total, p, b);
    return 0;
}
int orc_get_fields(const orc_ctx* c, double* u, double* v, double* w, double* p) {
    size_t b = sizeof(double) * c->N;
    if (u) memcpy(u, c->u, b);
    if (v) memcpy(v, c->v, b);
    if (w) memcpy(w, c->w, b);
    if (p) memcpy(p, c->p_total, b);
    return 0;
}
int orc_set_wall_gradients(orc_ctx* c, const double* gpx, const double* gpy) {
    size_t b = sizeof(double) * 2 * (size_t)c->npl;
    if (gpx) memcpy(c->gpx, gpx, b);
    if (gpy) memcpy(c->gpy, gpy, b);
    return 0;
}
int orc_get_wall_gradients(const orc_ctx* c, double* gpx, double* gpy) {
    size_t b = sizeof(double) * 2 * (size_t)c->npl;
    if (gpx) memcpy(gpx, c->gpx, b);
    if (gpy) memcpy(gpy, c->gpy, b);
    return 0;
}
int orc_set_controller(orc_ctx* c, double G, double integral, double prev, double gsum, int gcount) {
    c->pi.G = G; c->pi.err_int = integral; c->pi.err_prev = prev; c->pi.gsum = gsum; c->pi.gcount = gcount;
    c->pi.gavg = gcount > 0 ? gsum / (double)gcount : G;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Initial conditions                                                         */
/* ------------------------------------------------------------------------- */
static uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double u01(uint64_t* s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* initialize_fields_3d + apply_solenoidal_perturbations (DNS_pressure_BC_3D.f90:909-962,3219-3265),
 * generate_analytical_perturbations + scale_perturbations + enforce_wall_boundaries
 * (perturbation_analytical_simple.f90:24-155).  perturbation_kind 1 is the seeded
 * stream-function noise: psi = sum_{0<=mx,my<=K, (mx,my)!=(0,0)} a sin(kx x+ph1) sin(ky y+ph2) (1-z^2),
 * a = 1/(mx^2+my^2), phases 2*pi*u01 from splitmix64(seed) drawn (ph1, ph2) per mode in
 * mx-major order; u' = -dpsi/dy, v' = dpsi/dx, w' = 0, scaled to max |u'| = amplitude. */
int orc_make_initial_field(const orc_params* p, double* u, double* v, double* w) {
    int nx = p->nx, ny = p->ny, nz = p->nz;
    size_t npl = (size_t)nx * ny, N = npl * nz;
    double* z = dalloc(nz); double* wz = dalloc(nz);
    orc_lgl(nz, z, wz);
    for (int i = 0; i < nx; i++) for (int j = 0; j < ny; j++) for (int k = 0; k < nz; k++) {
        size_t n = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
        u[n] = p->u_centerline * (1.0 - z[k] * z[k]);
        v[n] = 0.0; w[n] = 0.0;
    }
    if (p->enable_perturbations) {
        double* up = dalloc(N); double* vp = dalloc(N);
        if (p->perturbation_kind == 0) {
            double kx = 2.0 * PI * (double)p->analytical_mode_x / p->xlen;
            double ky = 2.0 * PI * (double)p->analytical_mode_y / p->ylen;
            for (int k = 0; k < nz; k++) {
                double wd = 1.0 - z[k] * z[k];
                for (int j = 0; j < ny; j++) {
                    double y = (double)j * p->ylen / (double)ny;
                    for (int i = 0; i < nx; i++) {
                        double x = (double)i * p->xlen / (double)nx;
                        size_t n = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                        up[n] = -(ky * sin(kx * x) * cos(ky * y) * wd);
                        vp[n] = kx * cos(kx * x) * sin(ky * y) * wd;
                    }
                }
            }
        } else {
            int K = p->noise_kmax;
            uint64_t st = p->seed;
            double* u2 = dalloc(npl); double* v2 = dalloc(npl);
            for (int mx = 0; mx <= K; mx++) for (int my = 0; my <= K; my++) {
                if (mx == 0 && my == 0) continue;
                double ph1 = 2.0 * PI * u01(&st), ph2 = 2.0 * PI * u01(&st);
                double a = 1.0 / (double)(mx * mx + my * my);
                double kx = 2.0 * PI * (double)mx / p->xlen, ky = 2.0 * PI * (double)my / p->ylen;
                for (int j = 0; j < ny; j++) {
                    double y = (double)j * p->ylen / (double)ny;
                    double sy = sin(ky * y + ph2), cy = cos(ky * y + ph2);
                    for (int i = 0; i < nx; i++) {
                        double x = (double)i * p->xlen / (double)nx;
                        double sx = sin(kx * x + ph1), cx = cos(kx * x + ph1);
                        u2[i + (size_t)nx * j] += -(a * ky * sx * cy);
                        v2[i + (size_t)nx * j] += a * kx * cx * sy;
                    }
                }
            }
            for (int k = 0; k < nz; k++) {
                double wd = 1.0 - z[k] * z[k];
                for (size_t m = 0; m < npl; m++) { up[m + npl * k] = u2[m] * wd; vp[m + npl * k] = v2[m] * wd; }
            }
            free(u2); free(v2);
        }
        double cm = 0.0;
        for (size_t n = 0; n < N; n++) { double s = sqrt(up[n] * up[n] + vp[n] * vp[n] + 0.0 * 0.0); if (s > cm) cm = s; }
        if (cm > 1.0e-12) {
            double sf = p->perturbation_amplitude / cm;
            for (size_t n = 0; n < N; n++) { up[n] = up[n] * sf; vp[n] = vp[n] * sf; }
        }
        for (size_t m = 0; m < npl; m++) { up[m] = 0.0; vp[m] = 0.0; up[npl * (nz - 1) + m] = 0.0; vp[npl * (nz - 1) + m] = 0.0; }
        for (size_t n = 0; n < N; n++) { u[n] = u[n] + up[n]; v[n] = v[n] + vp[n]; w[n] = w[n] + 0.0; }
        free(up); free(vp);
    }
    for (size_t m = 0; m < npl; m++) {
        u[m] = 0.0; v[m] = 0.0; w[m] = 0.0;
        u[npl * (nz - 1) + m] = 0.0; v[npl * (nz - 1) + m] = 0.0; w[npl * (nz - 1) + m] = 0.0;
    }
    free(z); free(wz);
    return 0;
}

int orc_init_fields(orc_ctx* c) {
    orc_make_initial_field(&c->p, c->u, c->v, c->w);
    memset(c->p_total, 0, sizeof(double) * c->N);
    memset(c->gpx, 0, sizeof(double) * 2 * (size_t)c->npl);
    memset(c->gpy, 0, sizeof(double) * 2 * (size_t)c->npl);
    c->time = 0.0; c->step = 0;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* component entry points                                                     */
/* ------------------------------------------------------------------------- */
int orc_fft_r2c(int nx, int ny, int nplanes, const double* in, double* out) {
    if ((nx & 1) || (ny & 1)) return 1;
    plan2d p; plan2d_init(&p, nx, ny, 0);
    int m = nx > ny ? nx : ny;
    cpx* buf = (cpx*)malloc(sizeof(cpx) * m); cpx* work = (cpx*)malloc(sizeof(cpx) * m);
    for (int k = 0; k < nplanes; k++)
        r2c_plane(&p, in + (size_t)nx * ny * k, (cpx*)out + (size_t)(nx / 2 + 1) * ny * k, buf, work);
    free(buf); free(work); plan2d_free(&p);
    return 0;
}
int orc_fft_c2r(int nx, int ny, int nplanes, const double* in, double* out) {
    if ((nx & 1) || (ny & 1)) return 1;
    plan2d p; plan2d_init(&p, nx, ny, 0);
    int m = nx > ny ? nx : ny;
    cpx* buf = (cpx*)malloc(sizeof(cpx) * m); cpx* work = (cpx*)malloc(sizeof(cpx) * m);
    cpx* tmp = (cpx*)malloc(sizeof(cpx) * (size_t)(nx / 2 + 1) * ny);
    for (int k = 0; k < nplanes; k++)
        c2r_plane(&p, (const cpx*)in + (size_t)(nx / 2 + 1) * ny * k, out + (size_t)nx * ny * k, tmp, buf, work);
    free(buf); free(work); free(tmp); plan2d_free(&p);
    return 0;
}
