/* f90dns_debug.h -- inspection hooks of the channel-DNS C ABI, used by the parity tests.
 * They have no counterpart in the reference (whose stage data are program globals that a
 * debugger or a write statement would read); they copy device buffers to the host. */
#ifndef F90DNS_DEBUG_H
#define F90DNS_DEBUG_H
#include "f90dns.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Copy one resident spectral buffer (complex (nx/2+1, ny, nz), interleaved) to host `out`:
 * 0 u, 1 v, 2 w, 3 p, 4 u*, 5 v*, 6 w*, 7-9 dealiased S (x, y, z), 10 omega_x, 11 omega_y,
 * 12 phi, 13 D1 phi (the last two are written by dns_solve_pressure only). */
int dns_debug_spectral(dns_ctx* ctx, int which, double* out);
/* out = D1 in on one spectral field (host arrays), through the MFMA z-contraction. */
int dns_debug_zderiv(dns_ctx* ctx, const double* spec_in, double* spec_out);
/* Run nsteps eager fused steps with HIP events between the kernel launches on the context
 * stream; out[i] = mean duration in ms of kernel i (7 slots, names from
 * dns_profile_kernel_name).  The fused step runs the flow controller inside the Helmholtz
 * launch and mode (0,0) inside the Poisson launch, so those two slots read ~0.  Advances the
 * simulation like dns_step. */
int dns_profile_step(dns_ctx* ctx, int nsteps, double* out);
const char* dns_profile_kernel_name(int i);
/* Lane layout of the FP64 MFMA used by the contractions: 4096 single products with one-hot
 * operands (A at lane b/64, B at lane b%64); out[b*256 + lane*4 + q] (4096*256 doubles). */
int dns_debug_mfma_probe(double* out);
/* The same probe for v_mfma_f32_16x16x4f32 (the mixed-precision contraction). */
int dns_debug_mfma_probe_f32(double* out);
/* The initial fields of initialize_fields_3d for p, generated point by point on the host
 * (host only, no device needed); dns_init_fields assembles the same fields on the device. */
int dns_debug_initial_field(const dns_params* p, double* u, double* v, double* w);
/* One operator of the per-mode solves for (nz, dt, Re, theta, D1 variant), host only, row-major
 * (which: 0 Helmholtz forward, 1 back, 2 c1, 3 c2, 4 eigenvalues, 5 Poisson forward, 6 back,
 * 7 D1 back, 8 eigenvalues, 9 mode-(0,0) matrix, 10 D1, 11 LGL weights, 12 D2). */
int dns_debug_operator(int nz, double dt, double re, int cn, int exact, int which, double* out, int* rows, int* cols);
/* repeats launches of a plain copy kernel over n doubles (n even) on fresh buffers; mean ms
 * per launch in ms_out.  Calibrates the HBM byte counters of a profiler: each launch reads
 * and writes 8n bytes. */
int dns_debug_copy(size_t n, int repeats, double* ms_out);
/* max relative reconstruction error of the Helmholtz and Poisson eigen bases. */
int dns_debug_operator_error(dns_ctx* ctx, double* out2);
/* The mode space of the per-mode solves for an nx x ny plane where nz puts them on their
 * four-wave kernels (nz of about 98 and more; below that they walk the storage columns), host
 * only: *ncols real columns
 * (at most 2 (nx/2+1) ny); for column c, so[c] = its storage real column (re/im interleaved,
 * mode i + (nx/2+1) j), s2o[c] = -1 or a second storage column that receives s2s[c] times the
 * value (a mirrored mode's part, or 0 for the imaginary part of a self-conjugate mode).  Arrays
 * hold 2 (nx/2+1) ny entries.  No counterpart in the reference (a layout of this library). */
int dns_debug_zmap(int nx, int ny, int* ncols, int* so, int* s2o, double* s2s);

/* The wavenumbers of setup_spectral_3d (DNS_pressure_BC_3D.f90:780-796) as the library computes
 * and uploads them: xw(nx/2+1), yw(ny); host only. */
int dns_debug_wavenumbers(int nx, int ny, double xlen, double ylen, double* xw, double* yw);
/* The 2/3 dealias mask of apply_dealias_truncation (DNS_pressure_BC_3D.f90:1163-1193) as
 * (nx/2+1, ny) bytes, kx fastest, 1 = kept: which 0 = the table the two-pass transforms read,
 * 1 = the bounds and test the plane kernels apply; host only. */
int dns_debug_dealias_mask(int nx, int ny, int which, unsigned char* mask);
/* Counters of the last jacgs pressure solve (pressure_solver 1): total iterations over all
 * real systems, and systems that stopped at the iteration limit without converging. */
int dns_debug_jacgs_stats(dns_ctx* ctx, unsigned long long* iterations, int* nonconverged);
/* The device-resident wavenumber tables of a context, copied back: xw(nx/2+1), yw(ny). */
int dns_debug_context_wavenumbers(dns_ctx* ctx, double* xw, double* yw);

#ifdef __cplusplus
}
#endif
#endif
