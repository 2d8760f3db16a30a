/* f90dns_ensemble.h -- ragged ensembles of channel flows on one MI355X (gfx950).
 *
 * No counterpart in the reference, which runs one flow per process
 * (DNS_pressure_BC_3D.f90:286-361).  An ensemble holds members of different grids
 * (nx, ny, nz), Reynolds numbers, controller targets and initial perturbations; members of
 * equal grid, dt, domain, scheme flags and precision form a group.  Every stage of the step
 * is one launch per kernel class over all members (plane-size class for the transforms, row
 * blocks for the wall-normal kernels), and a step of the whole ensemble is replayed from one
 * captured HIP graph.  Each member is computed by the same arithmetic as a context of one
 * (include/f90dns.h), so its fields are bitwise those of the same run alone.
 *
 * Error convention as include/f90dns.h: 0 or a DNS_E* code, message from
 * dns_ensemble_last_error (NULL: last creation error).
 */
#ifndef F90DNS_ENSEMBLE_H
#define F90DNS_ENSEMBLE_H

#include <stddef.h>

#include "f90dns.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dns_ensemble dns_ensemble;

int dns_ensemble_create(const dns_params* members, int nmembers, dns_ensemble** out);
int dns_ensemble_destroy(dns_ensemble* e);
const char* dns_ensemble_last_error(const dns_ensemble* e);
int dns_ensemble_members(const dns_ensemble* e);
int dns_ensemble_groups(const dns_ensemble* e);
size_t dns_ensemble_device_bytes(const dns_ensemble* e);
void* dns_ensemble_stream(dns_ensemble* e);
int dns_ensemble_member_shape(const dns_ensemble* e, int member, int* nx, int* ny, int* nz);

/* initial fields of every member (dns_init_fields per member) */
int dns_ensemble_init_fields(dns_ensemble* e);
int dns_ensemble_set_member_fields(dns_ensemble* e, int member, const double* u, const double* v, const double* w,
                                   const double* p);
int dns_ensemble_get_member_fields(dns_ensemble* e, int member, double* u, double* v, double* w, double* p);
int dns_ensemble_member_diagnostics(dns_ensemble* e, int member, dns_diag* d);

/* nsteps steps of every member; use_graph: replay the captured step.  Asynchronous. */
int dns_ensemble_step(dns_ensemble* e, int nsteps, int use_graph);
int dns_ensemble_synchronize(dns_ensemble* e);
/* min, p50, p99, mean device time per ensemble step (ms) of the last dns_ensemble_step */
int dns_ensemble_step_timing(dns_ensemble* e, double* out4);

/* Launch-shape tuning: lanes = streams the step is split over (1 to 4; groups are dealt to the
 * least loaded lane, largest first); interleave = z-kernel
 * tile order; promote_fp64[4] / promote_fp32[4] = levels of promotion (below) for the plane
 * kernels of size classes 0..3 (NULL = none; classes 4..6 unpromoted). */
int dns_ensemble_set_tuning(dns_ensemble* e, int lanes, int interleave, const int* promote_fp64,
                            const int* promote_fp32);
/* Plane-transform size classes, as (threads per plane, largest (nx/2)*ny): 0 (64, 512),
 * 1 (256, 2048), 2 (512, 4096), 3 (1024, 8192), 4 (1024, 18432; FP32 only), 5 (576, 4608),
 * 6 (768, 6144).  A plane runs in the smallest class that holds it, in the order
 * 0, 1, 2, 5, 6, 3, 4; levels[c] (c < nclasses <= DNS_ENSEMBLE_SIZE_CLASSES) moves the planes of
 * class c that many places up this order (past the end: as far as it goes), for the FP64
 * (fp32 = 0) or FP32 (fp32 = 1) transforms; other classes are left unpromoted.  Results do not
 * change.  Promoting every class present to one class gives one c2r and one r2c launch per
 * step (with set_tile_waves 2, 2, 2 and one lane, one launch per stage for nz of 33 to 97).
 * No counterpart in the reference. */
enum { DNS_ENSEMBLE_SIZE_CLASSES = 7 };
int dns_ensemble_set_promotion(dns_ensemble* e, int fp32, const int* levels, int nclasses);
/* Launches of one step for the current members and tuning, per stage: counts[5] = vorticity
 * z-derivative, c2r, r2c, Helmholtz, Poisson (over both lanes). */
int dns_ensemble_launch_counts(dns_ensemble* e, int* counts);
/* Tile shapes of the wall-normal kernels: zwaves[3] = the least waves per column block of the
 * vorticity z-derivative, Helmholtz and Poisson launches (0 = one wave per row-block pair of the
 * group's operators, 2 or 4).  Groups whose operators have fewer row-block pairs run at that
 * width (the extra waves stage and store, and idle in the contractions), so groups of different
 * nz (e.g. 33 and 65 at two waves) share one launch per stage.  Results do not change with the
 * width.  No counterpart in the reference (a launch shape of this library). */
int dns_ensemble_set_tile_waves(dns_ensemble* e, const int* zwaves);
/* Autotuner: times steps_per_trial graph replays of each candidate tuning (coordinate
 * search), keeps the fastest, restores every member's fields and controller state, and
 * writes a JSON report of the trials (truncated to report_len). */
int dns_ensemble_autotune(dns_ensemble* e, int steps_per_trial, char* report, size_t report_len);

/* HBM bytes a context of nmembers members of shape p holds (dns_device_bytes after
 * dns_create_batch), for admission against a memory budget. */
int dns_estimate_bytes(const dns_params* p, int nmembers, size_t* bytes);

#ifdef __cplusplus
}
#endif
#endif
