// server.hip -- a device-resident serving loop over a ragged ensemble (include/f90dns_server.h).
//
// Jobs (a run's parameters and a step count) are queued on submission and admitted when a
// slot of their shape is free or a new group of slots fits the HBM budget.  Idle groups (whose
// jobs have all finished) keep their memory for jobs of their shape and give it back, longest
// idle first, when a new group needs it.  Every tick advances all running jobs by one step with
// one replay of the ensemble's captured step, so jobs that arrive while others run join the next
// step (continuous batching at step granularity).  Finished jobs are retired (diagnostics,
// optionally fields, copied out) and their slots reused.  A job that opens a new group is set up
// while the previous step runs.  A job whose fields go non-finite is evicted alone (its flow
// state is copied out beside every health_every-th step and read when that step is harvested).
// Ticks are pipelined one step deep: the host waits for step t-1 after queueing step t, except
// when membership changes, which needs an idle stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/f90dns_server.h"
#include "ensemble.h"

using namespace dns;

namespace {

struct Job {
    int id = 0;
    dns_params p{};
    int nsteps = 0, launched = 0, done = 0;
    int state = DNS_JOB_QUEUED;
    int keep = 0;
    int group = -1, slot = -1;
    size_t fixed_bytes = 0, slot_bytes = 0;   // HBM of a group of n slots: fixed + n * slot
    double t_submit = 0, t_admit = 0, t_finish = 0;
    std::vector<double> lat;                  // device ms of each of its steps
    std::vector<double> u, v, w, pr;          // final fields (keep_fields)
    dns_diag diag{};
    std::string why;
};

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

bool same_group(const dns_params& a, const dns_params& b) {
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dt == b.dt && a.xlen == b.xlen && a.ylen == b.ylen &&
           a.use_crank_nicolson == b.use_crank_nicolson && a.use_dealias_23_rule == b.use_dealias_23_rule &&
           a.use_weak_form_divergence == b.use_weak_form_divergence && a.precision == b.precision &&
           a.lgl_derivative == b.lgl_derivative;
}

struct Pending {
    hipEvent_t a, b;
    std::vector<int> jobs;     // running jobs of that step
    // health check copied out with the step: (job id, index into the server's pinned
    // FlowState buffer), read when the step is harvested
    std::vector<std::pair<int, int>> health;
};

}  // namespace

struct dns_server {
    Ensemble ens;
    size_t budget = 0;
    int slots = 4;
    bool graph = true;
    int health_every = 16;
    std::map<int, Job> jobs;
    std::deque<int> queue;
    std::vector<std::vector<int>> slot_job;   // [group][slot] -> job id or -1
    // [group]: the tick at which the group's last job retired, -1 while a job runs in it.  An
    // idle group keeps its memory for jobs of its shape until a new group needs the budget.
    std::vector<long long> idle_since;
    std::deque<Pending> inflight;
    std::vector<hipEvent_t> free_events;
    int next_id = 1;
    long long ticks = 0, job_steps = 0, changes = 0, captures = 0;
    double change_ms = 0, capture_ms = 0;   // host time of membership changes and re-captures
    double gp_steps = 0, device_ms = 0;
    size_t peak_bytes = 0;
    int max_running = 0;
    std::vector<double> all_lat;
    std::string err;
    // pinned host copies of running jobs' flow states for the health checks (two halves, so a
    // check's copies can be in flight while the previous check's are read)
    FlowState* hflow = nullptr;
    int hflow_cap = 0, hflow_half = 0;
    std::vector<int> evict;    // jobs found non-finite, evicted at the next membership change
};

static int sfail(dns_server* s, int code, const std::string& m) {
    s->err = m;
    return code;
}

static hipEvent_t take_event(dns_server* s) {
    if (!s->free_events.empty()) {
        hipEvent_t e = s->free_events.back();
        s->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// wait for the oldest in-flight step and account its latency to its jobs
static int harvest_one(dns_server* s) {
    Pending p = s->inflight.front();
    s->inflight.pop_front();
    if (hipEventSynchronize(p.b) != hipSuccess) return sfail(s, DNS_EDEVICE, "hipEventSynchronize failed");
    // the bulk integral is formed every step from mode (0,0) of u: a non-finite value flags a
    // job whose fields have gone non-finite
    for (const auto& h : p.health) {
        const FlowState& fs = s->hflow[h.second];
        if (!std::isfinite(fs.Ub) || !std::isfinite(fs.G)) s->evict.push_back(h.first);
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) != hipSuccess) return sfail(s, DNS_EDEVICE, "hipEventElapsedTime failed");
    s->device_ms += ms;
    for (int id : p.jobs) {
        auto it = s->jobs.find(id);
        if (it == s->jobs.end()) continue;
        it->second.lat.push_back(ms);
        it->second.done += 1;
        s->all_lat.push_back(ms);
    }
    s->free_events.push_back(p.a);
    s->free_events.push_back(p.b);
    return 0;
}

static int drain(dns_server* s) {
    while (!s->inflight.empty()) {
        int r = harvest_one(s);
        if (r) return r;
    }
    return s->ens.synchronize() ? sfail(s, DNS_EDEVICE, s->ens.err) : 0;
}

static int retire(dns_server* s, Job& j, int state, const std::string& why) {
    dns_ctx* c = s->ens.groups[j.group].c;
    if (state == DNS_JOB_DONE) {
        if (dns_member_diagnostics(c, j.slot, &j.diag)) return sfail(s, DNS_EDEVICE, dns_last_error(c));
        j.diag.step = j.done;
        j.diag.time = j.done * j.p.dt;
        if (j.keep) {
            const size_t n = size_t(c->N);
            j.u.resize(n); j.v.resize(n); j.w.resize(n); j.pr.resize(n);
            if (dns_get_member_fields(c, j.slot, j.u.data(), j.v.data(), j.w.data(), j.pr.data()))
                return sfail(s, DNS_EDEVICE, dns_last_error(c));
        }
    }
    j.state = state;
    j.why = why;
    j.t_finish = now_s();
    s->ens.set_active(j.group, j.slot, false);
    s->slot_job[j.group][j.slot] = -1;
    // a group with no running job stays allocated, idle, for the next job of its shape
    bool empty = true;
    for (int x : s->slot_job[j.group]) empty = empty && x < 0;
    if (empty) s->idle_since[j.group] = s->ticks;
    j.group = j.slot = -1;
    return 0;
}

static size_t idle_bytes(dns_server* s) {
    size_t b = 0;
    for (size_t g = 0; g < s->ens.groups.size(); g++)
        if (s->ens.groups[g].c && s->idle_since[g] >= 0) b += dns_device_bytes(s->ens.groups[g].c);
    return b;
}

// give back the memory of idle groups, the longest idle first, until `need` more bytes fit
// the budget (all of them for need = SIZE_MAX); the stream must be idle
static int release_idle(dns_server* s, size_t need) {
    for (;;) {
        if (need != SIZE_MAX && s->ens.bytes() + need <= s->budget) return 0;
        int g = -1;
        for (size_t gi = 0; gi < s->ens.groups.size(); gi++)
            if (s->ens.groups[gi].c && s->idle_since[gi] >= 0 && (g < 0 || s->idle_since[gi] < s->idle_since[g]))
                g = int(gi);
        if (g < 0) return 0;
        if (s->ens.remove_group(g)) return sfail(s, DNS_EDEVICE, "removing an idle group failed");
        s->idle_since[g] = -1;
    }
}

// evict the jobs that the asynchronous health checks found non-finite (stream idle)
static int evict_flagged(dns_server* s) {
    for (int id : s->evict) {
        auto it = s->jobs.find(id);
        if (it == s->jobs.end() || it->second.state != DNS_JOB_RUNNING) continue;
        int r = retire(s, it->second, DNS_JOB_FAILED, "non-finite fields (bulk velocity)");
        if (r) return r;
    }
    s->evict.clear();
    return 0;
}

// queue copies of the running jobs' flow states after the step just enqueued (read when the
// step is harvested: no wait here)
static int queue_health(dns_server* s, Pending& pd, const std::vector<int>& run) {
    const int need = int(run.size());
    if (2 * need > s->hflow_cap) {
        // grow only with the stream idle (a copy may still target the old buffer)
        int r = drain(s);
        if (r) return r;
        if (s->hflow) (void)hipHostFree(s->hflow);
        s->hflow = nullptr;
        s->hflow_cap = 2 * std::max(need, 64);
        if (hipHostMalloc(reinterpret_cast<void**>(&s->hflow), sizeof(FlowState) * s->hflow_cap) != hipSuccess) {
            s->hflow_cap = 0;
            return sfail(s, DNS_ENOMEM, "hipHostMalloc of the health-check buffer failed");
        }
    }
    const int base = s->hflow_half * (s->hflow_cap / 2);
    s->hflow_half ^= 1;
    for (int k = 0; k < need; k++) {
        const Job& j = s->jobs[run[k]];
        dns_ctx* c = s->ens.groups[j.group].c;
        if (hipMemcpyAsync(s->hflow + base + k, c->flow + j.slot, sizeof(FlowState), hipMemcpyDeviceToHost,
                           s->ens.stream()) != hipSuccess)
            return sfail(s, DNS_EDEVICE, "flow-state copy failed");
        pd.health.emplace_back(run[k], base + k);
    }
    return 0;
}

static bool fits(dns_server* s, const Job& j, int nslots) {
    return s->ens.bytes() + j.fixed_bytes + size_t(nslots) * j.slot_bytes <= s->budget;
}

// whether the queue holds a job that can be placed now (host bookkeeping only)
static bool admissible(dns_server* s) {
    const size_t idle = idle_bytes(s);
    for (int id : s->queue) {
        const Job& j = s->jobs[id];
        if (j.fixed_bytes + j.slot_bytes > s->budget) return true;   // to be rejected
        if (s->ens.bytes() - idle + j.fixed_bytes + j.slot_bytes <= s->budget) return true;
        for (size_t gi = 0; gi < s->ens.groups.size(); gi++) {
            dns_ctx* c = s->ens.groups[gi].c;
            if (!c || !same_group(c->p, j.p)) continue;
            for (int x : s->slot_job[gi])
                if (x < 0) return true;
        }
    }
    return false;
}

// place queued jobs (first fit, in submission order); returns whether membership changed
// fresh: place only jobs that open a new group within the budget as it stands, whose set-up
// touches no memory of a running group (so it may overlap the step in flight); otherwise the
// stream is idle and idle groups may be released for a new one
static int admit(dns_server* s, bool& changed, bool fresh) {
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Job& j = s->jobs[*it];
        int g = -1, slot = -1;
        for (size_t gi = 0; gi < s->ens.groups.size() && g < 0; gi++) {
            dns_ctx* c = s->ens.groups[gi].c;
            if (!c || !same_group(c->p, j.p)) continue;
            for (size_t k = 0; k < s->slot_job[gi].size(); k++)
                if (s->slot_job[gi][k] < 0) { g = int(gi); slot = int(k); break; }
        }
        if (g >= 0 && fresh) {
            ++it;   // a slot of a running group: placed once the stream is idle
            continue;
        }
        if (g < 0) {
            // a new group: as many slots as fit, up to the configured number
            if (!fresh && j.fixed_bytes + j.slot_bytes <= s->budget) {
                int r = release_idle(s, j.fixed_bytes + j.slot_bytes);
                if (r) return r;
            }
            int n = s->slots;
            while (n > 0 && !fits(s, j, n)) n--;
            if (n == 0 && fresh) {
                ++it;
                continue;
            }
            if (n == 0) {
                if (j.fixed_bytes + j.slot_bytes > s->budget) {
                    j.state = DNS_JOB_REJECTED;
                    j.why = "job does not fit the HBM budget";
                    it = s->queue.erase(it);
                    continue;
                }
                ++it;   // wait for memory
                continue;
            }
            std::vector<dns_params> ps(n, j.p);
            g = s->ens.add_group(ps.data(), n, false);
            if (g < 0) {
                j.state = DNS_JOB_REJECTED;
                j.why = s->ens.err;
                it = s->queue.erase(it);
                continue;
            }
            if (int(s->slot_job.size()) <= g) s->slot_job.resize(g + 1);
            if (int(s->idle_since.size()) <= g) s->idle_since.resize(g + 1, -1);
            s->slot_job[g].assign(n, -1);
            slot = 0;
            s->peak_bytes = std::max(s->peak_bytes, s->ens.bytes());
        }
        dns_ctx* c = s->ens.groups[g].c;
        if (dns_set_member_params(c, slot, &j.p) || dns_init_member_fields(c, slot))
            return sfail(s, DNS_EDEVICE, dns_last_error(c));
        s->ens.set_active(g, slot, true);
        s->slot_job[g][slot] = j.id;
        s->idle_since[g] = -1;
        j.group = g; j.slot = slot;
        j.state = DNS_JOB_RUNNING;
        j.t_admit = now_s();
        changed = true;
        it = s->queue.erase(it);
    }
    return 0;
}

extern "C" {

const char* dns_server_last_error(const dns_server* s) { return s ? s->err.c_str() : "null server"; }

int dns_server_create(size_t hbm_budget_bytes, int slots_per_group, int use_graph, dns_server** out) {
    if (!out || slots_per_group < 1 || hbm_budget_bytes == 0) return DNS_EINVAL;
    *out = nullptr;
    dns_server* s = new dns_server();
    int r = s->ens.init();
    if (r) { delete s; return r; }
    // membership changes too often to autotune; the step split over two streams (groups dealt to
    // the lighter lane) served the mixed sweep 3 % faster than one stream
    s->ens.tune.lanes = 2;
    s->budget = hbm_budget_bytes;
    s->slots = slots_per_group;
    s->graph = use_graph != 0;
    *out = s;
    return 0;
}

int dns_server_destroy(dns_server* s) {
    if (!s) return 0;
    (void)drain(s);
    if (s->hflow) (void)hipHostFree(s->hflow);
    for (auto e : s->free_events) (void)hipEventDestroy(e);
    delete s;
    return 0;
}

int dns_server_submit(dns_server* s, const dns_params* p, int nsteps, int keep_fields, int* id) {
    if (!s || !p || nsteps < 1) return DNS_EINVAL;
    if (p->convection_scheme != 0 || p->pressure_solver != 0)
        return sfail(s, DNS_EUNSUPPORTED, "the convection and pressure-solver variants are not served");
    size_t b1 = 0, b2 = 0;
    std::string e;
    if (!ctx_bytes_for(*p, 1, b1, e) || !ctx_bytes_for(*p, 2, b2, e)) return sfail(s, DNS_ESOLVER, e);
    Job j;
    j.slot_bytes = b2 - b1;
    j.fixed_bytes = b1 - j.slot_bytes;
    j.id = s->next_id++;
    j.p = *p;
    j.nsteps = nsteps;
    j.keep = keep_fields ? 1 : 0;
    j.t_submit = now_s();
    s->jobs[j.id] = j;
    s->queue.push_back(j.id);
    if (id) *id = j.id;
    return 0;
}

int dns_server_tick(dns_server* s, int* running) {
    if (!s) return DNS_EINVAL;
    int r;
    // jobs whose last step is queued, and admissible queued jobs, change the membership
    bool finish = false;
    for (auto& kv : s->jobs)
        if (kv.second.state == DNS_JOB_RUNNING && kv.second.launched >= kv.second.nsteps) finish = true;
    const bool health = s->health_every > 0 && s->ticks > 0 && s->ticks % s->health_every == 0;
    bool maybe_admit = !s->queue.empty() && admissible(s);
    if (finish || !s->evict.empty() || maybe_admit) {
        const double t0 = now_s();
        bool changed = false;
        // new groups are allocated and their jobs set up while the previous step runs
        if (maybe_admit && (r = admit(s, changed, true))) return r;
        if ((r = drain(s))) return r;
        if ((r = evict_flagged(s))) return r;
        for (auto& kv : s->jobs) {
            Job& j = kv.second;
            if (j.state == DNS_JOB_RUNNING && j.done >= j.nsteps && (r = retire(s, j, DNS_JOB_DONE, ""))) return r;
        }
        if ((r = admit(s, changed, false))) return r;
        s->changes += 1;
        s->change_ms += (now_s() - t0) * 1e3;
    }
    std::vector<int> run;
    for (auto& kv : s->jobs)
        if (kv.second.state == DNS_JOB_RUNNING && kv.second.launched < kv.second.nsteps) run.push_back(kv.first);
    if (running) *running = int(run.size());
    if (run.empty()) return 0;
    s->max_running = std::max(s->max_running, int(run.size()));
    Pending pd;
    pd.a = take_event(s);
    pd.b = take_event(s);
    if (!pd.a || !pd.b) return sfail(s, DNS_EDEVICE, "hipEventCreate failed");
    pd.jobs = run;
    // a changed membership rebuilds the tables and re-captures the step before the tick's
    // start event, so the tick's device time holds only the step
    if (s->ens.stale()) {
        const double t1 = now_s();
        if ((r = s->ens.prepare(s->graph))) return sfail(s, r, s->ens.err);
        s->captures += 1;
        s->capture_ms += (now_s() - t1) * 1e3;
    }
    if (hipEventRecord(pd.a, s->ens.stream()) != hipSuccess) return sfail(s, DNS_EDEVICE, "hipEventRecord failed");
    if ((r = s->ens.step(1, s->graph))) return sfail(s, r, s->ens.err);
    if (health && (r = queue_health(s, pd, run))) return r;
    if (hipEventRecord(pd.b, s->ens.stream()) != hipSuccess) return sfail(s, DNS_EDEVICE, "hipEventRecord failed");
    for (int id : run) {
        Job& j = s->jobs[id];
        j.launched += 1;
        s->gp_steps += double(j.p.nx) * j.p.ny * j.p.nz;
    }
    s->job_steps += long(run.size());
    s->ticks += 1;
    s->inflight.push_back(pd);
    while (s->inflight.size() > 1)
        if ((r = harvest_one(s))) return r;
    return 0;
}

int dns_server_drain(dns_server* s) {
    if (!s) return DNS_EINVAL;
    int r;
    for (;;) {
        int running = 0;
        if ((r = dns_server_tick(s, &running))) return r;
        bool left = !s->queue.empty();
        for (auto& kv : s->jobs) left = left || kv.second.state == DNS_JOB_RUNNING;
        if (!left) break;
        if (running == 0 && s->queue.empty()) {
            // every running job has all its steps queued: retire them (the non-finite ones failed)
            if ((r = drain(s)) || (r = evict_flagged(s))) return r;
            for (auto& kv : s->jobs)
                if (kv.second.state == DNS_JOB_RUNNING && (r = retire(s, kv.second, DNS_JOB_DONE, ""))) return r;
        }
    }
    return drain(s);
}

int dns_server_release_idle(dns_server* s) {
    if (!s) return DNS_EINVAL;
    int r = drain(s);
    return r ? r : release_idle(s, SIZE_MAX);
}

int dns_server_job_status(dns_server* s, int id, int* state, int* steps_done) {
    if (!s) return DNS_EINVAL;
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(s, DNS_EINVAL, "unknown job");
    if (state) *state = it->second.state;
    if (steps_done) *steps_done = it->second.done;
    return 0;
}

const char* dns_server_job_message(dns_server* s, int id) {
    if (!s) return "";
    auto it = s->jobs.find(id);
    return it == s->jobs.end() ? "unknown job" : it->second.why.c_str();
}

int dns_server_job_result(dns_server* s, int id, double* u, double* v, double* w, double* p, dns_diag* d) {
    if (!s) return DNS_EINVAL;
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(s, DNS_EINVAL, "unknown job");
    Job& j = it->second;
    if (j.state != DNS_JOB_DONE) return sfail(s, DNS_EINVAL, "job has not finished");
    if ((u || v || w || p) && !j.keep) return sfail(s, DNS_EINVAL, "job was submitted without keep_fields");
    const size_t n = j.u.size();
    if (u) std::memcpy(u, j.u.data(), n * sizeof(double));
    if (v) std::memcpy(v, j.v.data(), n * sizeof(double));
    if (w) std::memcpy(w, j.w.data(), n * sizeof(double));
    if (p) std::memcpy(p, j.pr.data(), n * sizeof(double));
    if (d) *d = j.diag;
    return 0;
}

int dns_server_job_times(dns_server* s, int id, double* out5) {
    if (!s || !out5) return DNS_EINVAL;
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(s, DNS_EINVAL, "unknown job");
    const Job& j = it->second;
    out5[0] = j.t_admit > 0 ? j.t_admit - j.t_submit : -1.0;
    out5[1] = j.t_finish > 0 ? j.t_finish - j.t_submit : -1.0;
    std::vector<double> t = j.lat;
    std::sort(t.begin(), t.end());
    auto pct = [&](double q) {
        if (t.empty()) return 0.0;
        const double pos = q * (t.size() - 1);
        const size_t lo = size_t(std::floor(pos)), hi = std::min(t.size() - 1, lo + 1);
        return t[lo] + (pos - lo) * (t[hi] - t[lo]);
    };
    out5[2] = pct(0.5);
    out5[3] = pct(0.99);
    out5[4] = double(t.size());
    return 0;
}

int dns_server_get_stats(dns_server* s, dns_server_stats* st) {
    if (!s || !st) return DNS_EINVAL;
    std::memset(st, 0, sizeof(*st));
    st->ticks = s->ticks;
    st->job_steps = s->job_steps;
    st->grid_point_steps = s->gp_steps;
    st->device_ms = s->device_ms;
    st->peak_bytes = s->peak_bytes;
    st->bytes = s->ens.bytes();
    st->max_running = s->max_running;
    st->changes = s->changes;
    st->captures = s->captures;
    st->change_ms = s->change_ms;
    st->capture_ms = s->capture_ms;
    for (auto& kv : s->jobs) {
        switch (kv.second.state) {
            case DNS_JOB_QUEUED: st->queued++; break;
            case DNS_JOB_RUNNING: st->running++; break;
            case DNS_JOB_DONE: st->done++; break;
            case DNS_JOB_FAILED: st->failed++; break;
            default: st->rejected++; break;
        }
    }
    std::vector<double> t = s->all_lat;
    std::sort(t.begin(), t.end());
    auto pct = [&](double q) {
        if (t.empty()) return 0.0;
        const double pos = q * (t.size() - 1);
        const size_t lo = size_t(std::floor(pos)), hi = std::min(t.size() - 1, lo + 1);
        return t[lo] + (pos - lo) * (t[hi] - t[lo]);
    };
    st->step_p50_ms = pct(0.5);
    st->step_p99_ms = pct(0.99);
    return 0;
}

int dns_server_pending(const dns_server* s) {
    if (!s) return 0;
    int n = int(s->queue.size());
    for (auto& kv : s->jobs) n += kv.second.state == DNS_JOB_RUNNING;
    return n;
}

int dns_server_set_health_interval(dns_server* s, int ticks) {
    if (!s || ticks < 0) return DNS_EINVAL;
    s->health_every = ticks;
    return 0;
}

}  // extern "C"
