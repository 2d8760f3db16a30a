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
 * stream; out[i] = mean duration in ms of kernel i (7 kernels, names from
 * dns_profile_kernel_name).  Advances the simulation like dns_step. */
int dns_profile_step(dns_ctx* ctx, int nsteps, double* out);
const char* dns_profile_kernel_name(int i);
/* Lane layout of the FP64 MFMA used by the contractions: 4096 single products with one-hot
 * operands (A at lane b/64, B at lane b%64); out[b*256 + lane*4 + q] (4096*256 doubles). */
int dns_debug_mfma_probe(double* out);
/* The same probe for v_mfma_f32_16x16x4f32 (the mixed-precision contraction). */
int dns_debug_mfma_probe_f32(double* out);
/* max relative reconstruction error of the Helmholtz and Poisson eigen bases. */
int dns_debug_operator_error(dns_ctx* ctx, double* out2);

#ifdef __cplusplus
}
#endif
#endif
