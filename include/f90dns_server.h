/* f90dns_server.h -- serving loop for parameter-sweep channel runs on one MI355X (gfx950).
 *
 * No counterpart in the reference, which runs one flow per process
 * (DNS_pressure_BC_3D.f90:286-361).  Jobs (dns_params + a step count) are queued, admitted
 * by HBM budget into slots of shape groups of a ragged ensemble (include/f90dns_ensemble.h),
 * kept resident on the GPU between steps and advanced together: every tick runs one step
 * of every running job, so a job submitted while others run joins at the next tick
 * (continuous batching).  A finished job gives its slot back (a group left without jobs
 * stays allocated for jobs of its shape until a new group needs the budget); a job whose
 * fields go non-finite is evicted alone.  A job's result is bitwise the result of the same run alone.
 */
#ifndef F90DNS_SERVER_H
#define F90DNS_SERVER_H

#include <stddef.h>

#include "f90dns.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { DNS_JOB_QUEUED = 0, DNS_JOB_RUNNING = 1, DNS_JOB_DONE = 2, DNS_JOB_FAILED = 3, DNS_JOB_REJECTED = 4 };

typedef struct dns_server_stats {
    long long ticks, job_steps;
    double grid_point_steps;     /* sum over job steps of nx*ny*nz */
    double device_ms;            /* device time of all ticks */
    double step_p50_ms, step_p99_ms;   /* per job-step latency (device time of its tick) */
    size_t bytes, peak_bytes;    /* HBM held by the groups (idle ones included) now / at most */
    int max_running;
    int queued, running, done, failed, rejected;
    /* host time of the serving loop: membership changes (drain, retire, admit) and re-captures
       of the step graph after them, during which the device is idle */
    long long changes, captures;
    double change_ms, capture_ms;
} dns_server_stats;

typedef struct dns_server dns_server;

/* slots_per_group: slots a new shape group gets (fewer if the budget is short).  Each tick's
 * step is split over two streams (include/f90dns_ensemble.h, lanes). */
int dns_server_create(size_t hbm_budget_bytes, int slots_per_group, int use_graph, dns_server** out);
int dns_server_destroy(dns_server* s);
const char* dns_server_last_error(const dns_server* s);
/* queue a job; keep_fields: keep the final u, v, w, p for dns_server_job_result */
int dns_server_submit(dns_server* s, const dns_params* p, int nsteps, int keep_fields, int* job_id);
/* retire finished jobs, admit queued ones, run one step of every running job */
int dns_server_tick(dns_server* s, int* running);
/* tick until no job is queued or running */
int dns_server_drain(dns_server* s);
int dns_server_job_status(dns_server* s, int job_id, int* state, int* steps_done);
const char* dns_server_job_message(dns_server* s, int job_id);
int dns_server_job_result(dns_server* s, int job_id, double* u, double* v, double* w, double* p, dns_diag* d);
/* queue wait (s), submit-to-finish (s), p50 and p99 of its step latency (ms), steps timed */
int dns_server_job_times(dns_server* s, int job_id, double* out5);
int dns_server_get_stats(dns_server* s, dns_server_stats* st);
/* jobs queued or running */
int dns_server_pending(const dns_server* s);
/* ticks between non-finite checks (0 = never; default 16) */
int dns_server_set_health_interval(dns_server* s, int ticks);
/* free the HBM of idle groups (shape groups whose jobs have all finished).  Idle groups are
   otherwise kept for jobs of their shape and released, longest idle first, only when a new
   group needs the budget. */
int dns_server_release_idle(dns_server* s);

#ifdef __cplusplus
}
#endif
#endif
