// ensemble.hip -- ragged ensembles (ensemble.h) and their C ABI (include/f90dns_ensemble.h).
#include "ensemble.h"

#include "contract.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <sstream>

#include "../../include/f90dns_ensemble.h"
#include "setup.h"

namespace dns {

#define ECHECK(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                             \
            return DNS_EDEVICE;                                                                  \
        }                                                                                        \
    } while (0)

bool ctx_bytes_for(const dns_params& p, int B, size_t& bytes, std::string& err) {
    Operators o;
    if (!build_operators(p.nz, p.dt, p.re, p.use_crank_nicolson ? 0.5 : 1.0, p.use_crank_nicolson != 0,
                         p.lgl_derivative != 0, p.use_weak_form_divergence != 0, o, err))
        return false;
    const size_t nxh = p.nx / 2 + 1, H2 = 2 * nxh * p.ny, nz = p.nz, m = nz - 2;
    const size_t NS = H2 * nz, N = size_t(p.nx) * p.ny * nz;
    auto d = [](size_t n) { return std::max<size_t>(n, 1) * sizeof(double); };
    size_t b = 0;
    b += d(2 * p.nx) + std::max<size_t>(2 * p.nx, 1) * sizeof(float);   // x twiddles
    b += d(2 * p.ny) + std::max<size_t>(2 * p.ny, 1) * sizeof(float);   // y twiddles
    b += d(p.nx / 2 + 1) + d(p.ny) + std::max<size_t>(nxh * p.ny, 1);  // wavenumbers, mask
    b += d(nz) + d(nz * nz) * 2 + d(o.zD1.data.size());
    if (p.precision == DNS_PREC_MIXED_F32D) b += std::max<size_t>(o.zD1.data.size(), 1) * sizeof(float);
    b += d(o.hF.data.size() * B) + d(o.hB.data.size()) + 3 * d(m * B) + d(B);
    b += d(o.pF.data.size()) + d(o.pB.data.size()) + d(o.pDB.data.size()) + d(nz);
    const ZMap zm = zmap_make(p.nx, p.ny, ops_wide(o));
    b += size_t(2) * zm.P * sizeof(int4) + std::max<size_t>(zmap_blocks(zm), 1) * sizeof(int);   // mode-space tables
    b += 15 * d(NS * B) + 6 * d(N * B);
    b += std::max<size_t>(B, 1) * (sizeof(FlowState) + sizeof(FlowMember));
    b += d(512 * 5) + d(5);
    if (p.pressure_solver == 1)   // jacgs: stiffness and the two solve counters
        b += d(nz * nz) + sizeof(unsigned long long) + sizeof(int);
    FFTGeom g;
    fft_geom(p.nx, p.ny, g);
    if (g.cls64 < 0) b += d(NS);   // two-pass scratch (plane beyond the FP64 LDS transform)
    bytes = b;
    return true;
}

Ensemble::~Ensemble() {
    for (hipStream_t s : sl)
        if (s) (void)hipStreamSynchronize(s);
    drop_graph();
    for (auto e : ev) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    for (hipEvent_t e : join)
        if (e) (void)hipEventDestroy(e);
    void* bufs[] = {d_c2r, d_r2c, d_z, d_h, d_p, d_tiles};
    for (void* q : bufs)
        if (q) (void)hipFree(q);
    for (auto& g : groups)
        if (g.c) dns_destroy(g.c);
    for (hipStream_t s : sl)
        if (s) (void)hipStreamDestroy(s);
}

int Ensemble::init() {
    std::string why;
    if (!dns_check_gfx950(why)) { err = why; return DNS_EDEVICE; }
    for (int l = 0; l < ENS_LANES; l++) {
        ECHECK(hipStreamCreateWithFlags(&sl[l], hipStreamNonBlocking));
        ECHECK(hipEventCreateWithFlags(&join[l], hipEventDisableTiming));
    }
    s0 = sl[0];
    ECHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    ECHECK(fft_prepare_all());
    return 0;
}

int Ensemble::add_group(const dns_params* params, int n, bool active) {
    dns_ctx* c = nullptr;
    const int rc = dns_create_batch(params, n, &c);
    if (rc) {
        err = dns_last_error(nullptr);
        return -1;
    }
    Group g;
    g.c = c;
    g.active.assign(n, active ? 1 : 0);
    groups.push_back(g);
    dirty = true;
    return int(groups.size()) - 1;
}

int Ensemble::remove_group(int g) {
    if (g < 0 || g >= int(groups.size()) || !groups[g].c) return DNS_EINVAL;
    for (hipStream_t s : sl) ECHECK(hipStreamSynchronize(s));
    drop_graph();
    dns_destroy(groups[g].c);
    groups[g].c = nullptr;
    groups[g].active.clear();
    dirty = true;
    return 0;
}

int Ensemble::set_active(int g, int slot, bool on) {
    if (g < 0 || g >= int(groups.size()) || !groups[g].c || slot < 0 || slot >= int(groups[g].active.size()))
        return DNS_EINVAL;
    if (bool(groups[g].active[slot]) != on) {
        groups[g].active[slot] = on ? 1 : 0;
        dirty = true;
    }
    return 0;
}

int Ensemble::live_slots() const {
    int n = 0;
    for (auto& g : groups)
        for (char a : g.active) n += a;
    return n;
}

size_t Ensemble::bytes() const {
    size_t b = 0;
    for (auto& g : groups)
        if (g.c) b += dns_device_bytes(g.c);
    return b;
}

int Ensemble::drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    return 0;
}

template <class T>
static hipError_t grow(T** p, size_t& cap, size_t n) {
    if (n <= cap && *p) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(n, 1) + n / 2;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), want * sizeof(T));
    if (e == hipSuccess) cap = want;
    return e;
}

int Ensemble::rebuild() {
    for (hipStream_t s : sl) ECHECK(hipStreamSynchronize(s));
    drop_graph();
    const int G = int(groups.size());
    std::vector<FFTArgs> hc2r(G), hr2c(G);
    std::vector<ZArgs> hz(G);
    std::vector<HelmArgs> hh(G);
    std::vector<PoisArgs> hp(G);
    for (int g = 0; g < G; g++) {
        std::memset(&hc2r[g], 0, sizeof(FFTArgs));
        std::memset(&hr2c[g], 0, sizeof(FFTArgs));
        dns_ctx* c = groups[g].c;
        if (!c) continue;
        hc2r[g] = ctx_c2r_nl_args(c);
        fft_fill(c->g, hc2r[g], c->mixed);
        hr2c[g] = ctx_r2c_cross_args(c);
        fft_fill(c->g, hr2c[g], c->mixed);
        hz[g] = ctx_vort_args(c);
        hh[g] = ctx_helm_args(c);
        hh[g].fl = ctx_flow_args(c);
        hh[g].flow_fused = 1;   // the flow controller runs in the Helmholtz launch
        hp[g] = ctx_pois_args(c, 1);   // mode (0,0) is solved in the Poisson launch
        hz[g].zwaves = tune.zwaves[0];
        hh[g].zwaves = tune.zwaves[1];
        hp[g].zwaves = tune.zwaves[2];
    }
    // lane assignment: balance an estimate of the work (points x (log2 plane + nz/8))
    std::vector<double> load(ENS_LANES, 0.0);
    std::vector<int> order;
    for (int g = 0; g < G; g++)
        if (groups[g].c) order.push_back(g);
    auto cost = [&](int g) {
        dns_ctx* c = groups[g].c;
        int act = 0;
        for (char a : groups[g].active) act += a;
        return double(act) * c->N * (std::log2(double(c->npl)) + c->nz / 8.0);
    };
    std::sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
    const int nlanes = std::max(1, std::min(tune.lanes, ENS_LANES));
    for (int g : order) {   // largest first, each to the least loaded lane
        int l = 0;
        for (int k = 1; k < nlanes; k++)
            if (load[k] < load[l]) l = k;
        groups[g].lane = l;
        load[l] += cost(g);
    }
    // tiles per (lane, kind, class)
    std::vector<Tile> tiles;
    launches.clear();
    for (int lane = 0; lane < nlanes; lane++) {
        for (int kind = 0; kind < K_COUNT; kind++) {
            // class key -> groups
            std::map<std::pair<int, int>, std::vector<int>> cls;
            for (int g = 0; g < G; g++) {
                dns_ctx* c = groups[g].c;
                if (!c || groups[g].lane != lane) continue;
                bool any = false;
                for (char a : groups[g].active) any = any || a;
                if (!any) continue;
                int key = 0, f32 = c->mixed ? 1 : 0;
                if (kind == K_C2R || kind == K_R2C) {
                    const int base = c->mixed ? c->g.cls32 : c->g.cls64;
                    int lvl = tune.promote[f32][base];
                    while (lvl > 0 && fft_class_promote(base, lvl, c->mixed) < 0) lvl--;
                    key = fft_class_promote(base, lvl, c->mixed);
                } else if (kind == K_ZVORT) {
                    key = zd_variant(hz[g]); f32 = hz[g].d1_f32 ? 1 : 0;
                } else if (kind == K_POIS) {
                    key = pois_variant(hp[g]); f32 = 0;
                } else if (kind == K_HELM) {
                    key = helm_variant(hh[g]); f32 = hh[g].s_f32 ? 1 : 0;
                } else {
                    f32 = 0;
                }
                cls[{f32, key}].push_back(g);
            }
            for (auto& kv : cls) {
                Launch L{};
                L.kind = kind; L.f32 = kv.first.first; L.cls = kv.first.second;
                L.off = int(tiles.size()); L.lane = lane; L.lds = 0; L.nf = 1;
                for (int g : kv.second) {
                    dns_ctx* c = groups[g].c;
                    const auto& act = groups[g].active;
                    const int B = int(act.size());
                    if (kind == K_C2R || kind == K_R2C) {
                        L.nf = kind == K_C2R ? 6 : 3;
                        L.lds = std::max(L.lds, plane_lds_bytes(c->g, c->mixed));
                        for (int z = 0; z < B; z++)
                            if (act[z])
                                for (int k = 1; k < c->nz - 1; k++) tiles.push_back(Tile{g, z * c->nz + k, z});
                    } else if (kind == K_HELM || kind == K_POIS || kind == K_ZVORT) {
                        L.lds = std::max(L.lds, kind == K_HELM   ? helm_lds_bytes(hh[g])
                                                : kind == K_POIS ? pois_lds_bytes(hp[g])
                                                                 : zderiv_lds_bytes(hz[g]));
                        // the Helmholtz and Poisson tiles walk the independent slots only
                        const int ncb = zmap_blocks(c->zm);
                        if (tune.interleave) {
                            for (int x = 0; x < ncb; x++)
                                for (int z = 0; z < B; z++)
                                    if (act[z]) tiles.push_back(Tile{g, x, z});
                        } else {
                            for (int z = 0; z < B; z++)
                                if (act[z])
                                    for (int x = 0; x < ncb; x++) tiles.push_back(Tile{g, x, z});
                        }
                    }
                }
                L.count = int(tiles.size()) - L.off;
                if (L.count > 0) launches.push_back(L);
            }
        }
    }
    if (G > 0) {
        ECHECK(grow(&d_c2r, cap_groups, G));
        size_t cg = 0;
        ECHECK(grow(&d_r2c, cg, G)); cg = 0;
        ECHECK(grow(&d_z, cg, G)); cg = 0;
        ECHECK(grow(&d_h, cg, G)); cg = 0;
        ECHECK(grow(&d_p, cg, G));
        ECHECK(hipMemcpy(d_c2r, hc2r.data(), G * sizeof(FFTArgs), hipMemcpyHostToDevice));
        ECHECK(hipMemcpy(d_r2c, hr2c.data(), G * sizeof(FFTArgs), hipMemcpyHostToDevice));
        ECHECK(hipMemcpy(d_z, hz.data(), G * sizeof(ZArgs), hipMemcpyHostToDevice));
        ECHECK(hipMemcpy(d_h, hh.data(), G * sizeof(HelmArgs), hipMemcpyHostToDevice));
        ECHECK(hipMemcpy(d_p, hp.data(), G * sizeof(PoisArgs), hipMemcpyHostToDevice));
    }
    if (!tiles.empty()) {
        ECHECK(grow(&d_tiles, cap_tiles, tiles.size()));
        ECHECK(hipMemcpy(d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
    }
    dirty = false;
    return 0;
}

int Ensemble::launch(const Launch& L, hipStream_t s) {
    const Tile* map = d_tiles + L.off;
    hipError_t e = hipSuccess;
    switch (L.kind) {
        case K_ZVORT: e = launch_zderiv_rag(L.cls, Z_VORT, L.f32 != 0, d_z, map, L.count, 1, L.lds, s); break;
        case K_C2R: e = launch_c2r_rag(L.cls, C2R_NL, L.f32 != 0, d_c2r, map, L.count, L.nf, L.lds, s); break;
        case K_R2C: e = launch_r2c_rag(L.cls, R2C_CROSS, L.f32 != 0, d_r2c, map, L.count, L.nf, L.lds, s); break;
        case K_HELM: e = launch_helmholtz_rag(L.cls, L.f32 != 0, d_h, map, L.count, L.lds, s); break;
        case K_POIS: e = launch_poisson_rag(L.cls, d_p, map, L.count, L.lds, s); break;
        default: e = hipErrorInvalidValue;
    }
    if (e != hipSuccess) {
        err = std::string("ragged launch: ") + hipGetErrorString(e);
        return DNS_EDEVICE;
    }
    return 0;
}

int Ensemble::enqueue_step() {
    if (dirty) {
        int r = rebuild();
        if (r) return r;
    }
    const int nlanes = std::max(1, std::min(tune.lanes, ENS_LANES));
    if (nlanes > 1) {
        ECHECK(hipEventRecord(fork, s0));
        for (int l = 1; l < nlanes; l++) ECHECK(hipStreamWaitEvent(sl[l], fork, 0));
    }
    // launches are stored lane by lane, each lane in stage order
    for (const Launch& L : launches) {
        int r = launch(L, sl[L.lane]);
        if (r) return r;
    }
    for (int l = 1; l < nlanes; l++) {
        ECHECK(hipEventRecord(join[l], sl[l]));
        ECHECK(hipStreamWaitEvent(s0, join[l], 0));
    }
    return 0;
}

int Ensemble::ensure_graph() {
    if (dirty) {
        int r = rebuild();
        if (r) return r;
    }
    if (gexec) return 0;
    ECHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    const int r = enqueue_step();
    hipGraph_t gph = nullptr;
    const hipError_t e = hipStreamEndCapture(s0, &gph);
    if (r) {
        if (gph) (void)hipGraphDestroy(gph);
        return r;
    }
    if (e != hipSuccess) {
        err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return DNS_EDEVICE;
    }
    graph = gph;
    ECHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    return 0;
}

void Ensemble::launch_counts(int* counts) const {
    for (int k = 0; k < K_COUNT; k++) counts[k] = 0;
    for (const Launch& L : launches) counts[L.kind] += 1;
}

int Ensemble::prepare(bool use_graph) {
    if (use_graph) return ensure_graph();
    return dirty ? rebuild() : 0;
}

int Ensemble::step(int n, bool use_graph) {
    if (n < 0) return DNS_EINVAL;
    int r;
    if (use_graph && (r = ensure_graph())) return r;
    if (!use_graph && dirty && (r = rebuild())) return r;
    while (int(ev.size()) < n + 1) {
        hipEvent_t e;
        ECHECK(hipEventCreate(&e));
        ev.push_back(e);
    }
    ECHECK(hipEventRecord(ev[0], s0));
    for (int s = 0; s < n; s++) {
        if (use_graph) {
            ECHECK(hipGraphLaunch(gexec, s0));
        } else if ((r = enqueue_step())) {
            return r;
        }
        for (auto& g : groups)
            if (g.c && (r = ctx_after_correction(g.c, s0))) return r;
        ECHECK(hipEventRecord(ev[s + 1], s0));
        for (auto& g : groups)
            if (g.c) dns_begin_step(g.c);
        steps_done++;
    }
    pending = n;
    return 0;
}

int Ensemble::synchronize() {
    ECHECK(hipStreamSynchronize(s0));
    return 0;
}

int Ensemble::timing(double* out4) {
    if (pending >= 0) {
        const int n = pending;
        pending = -1;
        step_ms.clear();
        if (n > 0) {
            ECHECK(hipEventSynchronize(ev[n]));
            for (int s = 0; s < n; s++) {
                float ms = 0.f;
                ECHECK(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
                step_ms.push_back(ms);
            }
        }
    }
    if (step_ms.empty()) { out4[0] = out4[1] = out4[2] = out4[3] = 0.0; return 0; }
    std::vector<double> t = step_ms;
    std::sort(t.begin(), t.end());
    auto pct = [&](double q) {
        const double pos = q * (t.size() - 1);
        const size_t lo = size_t(std::floor(pos)), hi = std::min(t.size() - 1, lo + 1);
        return t[lo] + (pos - lo) * (t[hi] - t[lo]);
    };
    double sum = 0;
    for (double x : t) sum += x;
    out4[0] = t.front(); out4[1] = pct(0.5); out4[2] = pct(0.99); out4[3] = sum / t.size();
    return 0;
}

int Ensemble::save_state(std::vector<void*>& bufs) {
    ECHECK(hipStreamSynchronize(s0));
    for (auto& g : groups) {
        if (!g.c) continue;
        dns_ctx* c = g.c;
        const size_t nb = c->NS * c->B * sizeof(double);
        double* fields[4] = {c->uh, c->vh, c->wh, c->ph};
        for (double* f : fields) {
            void* q = nullptr;
            ECHECK(hipMalloc(&q, nb));
            bufs.push_back(q);
            ECHECK(hipMemcpyAsync(q, f, nb, hipMemcpyDeviceToDevice, s0));
        }
        void* q = nullptr;
        ECHECK(hipMalloc(&q, sizeof(FlowState) * c->B));
        bufs.push_back(q);
        ECHECK(hipMemcpyAsync(q, c->flow, sizeof(FlowState) * c->B, hipMemcpyDeviceToDevice, s0));
    }
    ECHECK(hipStreamSynchronize(s0));
    return 0;
}

int Ensemble::restore_state(std::vector<void*>& bufs) {
    ECHECK(hipStreamSynchronize(s0));
    size_t i = 0;
    for (auto& g : groups) {
        if (!g.c) continue;
        dns_ctx* c = g.c;
        const size_t nb = c->NS * c->B * sizeof(double);
        double* fields[4] = {c->uh, c->vh, c->wh, c->ph};
        for (double* f : fields) ECHECK(hipMemcpyAsync(f, bufs[i++], nb, hipMemcpyDeviceToDevice, s0));
        ECHECK(hipMemcpyAsync(c->flow, bufs[i++], sizeof(FlowState) * c->B, hipMemcpyDeviceToDevice, s0));
    }
    ECHECK(hipStreamSynchronize(s0));
    for (void* q : bufs) (void)hipFree(q);
    bufs.clear();
    return 0;
}

int Ensemble::time_steps(int n, double& ms) {
    int r;
    if ((r = ensure_graph())) return r;
    for (int w = 0; w < 2; w++) ECHECK(hipGraphLaunch(gexec, s0));
    hipEvent_t a, b;
    ECHECK(hipEventCreate(&a));
    ECHECK(hipEventCreate(&b));
    ECHECK(hipEventRecord(a, s0));
    for (int s = 0; s < n; s++) ECHECK(hipGraphLaunch(gexec, s0));
    ECHECK(hipEventRecord(b, s0));
    ECHECK(hipEventSynchronize(b));
    float t = 0.f;
    ECHECK(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ms = double(t) / n;
    return 0;
}

// the tuning as JSON members (closed with a brace when closed)
static std::string tuning_json(const EnsTuning& t, bool closed) {
    std::ostringstream o;
    o << "{\"lanes\": " << t.lanes << ", \"interleave\": " << t.interleave;
    for (int f = 0; f < 2; f++) {
        o << (f ? ", \"promote_fp32\": [" : ", \"promote_fp64\": [");
        for (int c = 0; c < FFT_CLASSES; c++) o << (c ? "," : "") << t.promote[f][c];
        o << "]";
    }
    o << ", \"tile_waves\": [" << t.zwaves[0] << "," << t.zwaves[1] << "," << t.zwaves[2] << "]" << (closed ? "}" : "");
    return o.str();
}

int Ensemble::autotune(int steps, std::string& report) {
    if (steps < 1) steps = 1;
    std::vector<void*> saved;
    int r = save_state(saved);
    if (r) return r;
    std::vector<long long> gsteps;
    std::vector<double> gtime;
    for (auto& g : groups) {
        gsteps.push_back(g.c ? g.c->step : 0);
        gtime.push_back(g.c ? g.c->time : 0.0);
    }
    std::ostringstream rep;
    rep << "{\"trials\": [";
    bool first = true;
    auto trial = [&](const EnsTuning& t, double& ms) -> int {
        tune = t;
        dirty = true;
        int rr = time_steps(steps, ms);
        if (rr) return rr;
        rep << (first ? "" : ", ") << tuning_json(t, false) << ", \"ms_per_step\": " << ms << "}";
        first = false;
        return 0;
    };
    EnsTuning best = tune;
    double best_ms = 0;
    if ((r = trial(best, best_ms))) return r;
    // coordinate search: stream split, tile order, then each plane-kernel size class
    std::vector<EnsTuning> cand;
    {
        for (int l = 1; l <= ENS_LANES; l++)
            if (l != best.lanes) {
                EnsTuning t = best; t.lanes = l; cand.push_back(t);
            }
    }
    for (auto& t0 : cand) {
        double ms;
        if ((r = trial(t0, ms))) return r;
        if (ms < best_ms) { best_ms = ms; best = t0; }
    }
    {
        EnsTuning t = best; t.interleave = 1 - best.interleave;
        double ms;
        if ((r = trial(t, ms))) return r;
        if (ms < best_ms) { best_ms = ms; best = t; }
    }
    for (int f32 = 0; f32 < 2; f32++) {
        for (int c = 0; c < FFT_CLASSES; c++) {
            bool used = false;
            for (auto& g : groups)
                if (g.c && (g.c->mixed ? 1 : 0) == f32 && (f32 ? g.c->g.cls32 : g.c->g.cls64) == c) used = true;
            if (!used) continue;
            for (int lvl = 1; lvl <= 2; lvl++) {
                if (fft_class_promote(c, lvl, f32 != 0) < 0) break;
                EnsTuning t = best;
                t.promote[f32][c] = lvl;
                double ms;
                if ((r = trial(t, ms))) return r;
                if (ms < best_ms) { best_ms = ms; best = t; }
            }
        }
    }
    // tile shapes of the wall-normal kernels, one kernel kind at a time: widths (natural, two or
    // four waves) that change some group's variant (a group whose operators have one or two
    // row-block pairs)
    for (int k = 0; k < 3; k++) {
        for (int w : {0, 2, 4}) {
            if (w == best.zwaves[k]) continue;
            bool changes = false;
            for (auto& g : groups) {
                if (!g.c) continue;
                const int nat = k == 0 ? zd_variant(ctx_vort_args(g.c)) : k == 1 ? helm_variant(ctx_helm_args(g.c))
                                                                                : pois_variant(ctx_pois_args(g.c, 1));
                if (z_variant_widen(nat, w) != z_variant_widen(nat, best.zwaves[k])) changes = true;
            }
            if (!changes) continue;
            EnsTuning t = best;
            t.zwaves[k] = w;
            double ms;
            if ((r = trial(t, ms))) return r;
            if (ms < best_ms) { best_ms = ms; best = t; }
        }
    }
    // the stream split once more, for the tile shapes chosen
    for (int l = 1; l <= ENS_LANES; l++) {
        if (l == best.lanes) continue;
        EnsTuning t = best;
        t.lanes = l;
        double ms;
        if ((r = trial(t, ms))) return r;
        if (ms < best_ms) { best_ms = ms; best = t; }
    }
    tune = best;
    dirty = true;
    rep << "], \"chosen\": " << tuning_json(best, true) << ", \"ms_per_step\": " << best_ms << "}";
    report = rep.str();
    if ((r = restore_state(saved))) return r;
    for (size_t g = 0; g < groups.size(); g++)
        if (groups[g].c) { groups[g].c->step = int(gsteps[g]); groups[g].c->time = gtime[g]; }
    return 0;
}

}  // namespace dns

// ---- C ABI --------------------------------------------------------------------------------
using namespace dns;

struct dns_ensemble {
    Ensemble e;
    std::vector<std::pair<int, int>> where;   // member -> (group, slot)
};

namespace {
thread_local std::string g_ens_error;
}

static int efail(dns_ensemble* e, int code, const std::string& msg) {
    if (e) e->e.err = msg; else g_ens_error = msg;
    return code;
}

static bool same_group(const dns_params& a, const dns_params& b) {
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dt == b.dt && a.xlen == b.xlen && a.ylen == b.ylen &&
           a.use_crank_nicolson == b.use_crank_nicolson && a.use_dealias_23_rule == b.use_dealias_23_rule &&
           a.use_weak_form_divergence == b.use_weak_form_divergence && a.precision == b.precision &&
           a.lgl_derivative == b.lgl_derivative;
}

extern "C" {

const char* dns_ensemble_last_error(const dns_ensemble* e) { return e ? e->e.err.c_str() : g_ens_error.c_str(); }

int dns_ensemble_create(const dns_params* members, int n, dns_ensemble** out) {
    if (!members || n < 1 || !out) return efail(nullptr, DNS_EINVAL, "null argument or no members");
    *out = nullptr;
    for (int i = 0; i < n; i++)
        if (members[i].convection_scheme != 0 || members[i].pressure_solver != 0)
            return efail(nullptr, DNS_EUNSUPPORTED,
                         "member " + std::to_string(i) + ": the convection and pressure-solver variants run in "
                         "single contexts and batches only");
    dns_ensemble* e = new dns_ensemble();
    int r = e->e.init();
    if (r) { g_ens_error = e->e.err; delete e; return r; }
    // group members of equal shape, in order of first appearance
    std::vector<int> rep;
    std::vector<std::vector<int>> idx;
    for (int i = 0; i < n; i++) {
        size_t g = 0;
        while (g < rep.size() && !same_group(members[rep[g]], members[i])) g++;
        if (g == rep.size()) { rep.push_back(i); idx.emplace_back(); }
        idx[g].push_back(i);
    }
    e->where.assign(n, {-1, -1});
    for (size_t g = 0; g < rep.size(); g++) {
        std::vector<dns_params> ps;
        for (int i : idx[g]) ps.push_back(members[i]);
        const int gi = e->e.add_group(ps.data(), int(ps.size()), true);
        if (gi < 0) {
            g_ens_error = "member " + std::to_string(idx[g][0]) + ": " + e->e.err;
            delete e;
            return DNS_EINVAL;
        }
        for (size_t s = 0; s < idx[g].size(); s++) e->where[idx[g][s]] = {gi, int(s)};
    }
    *out = e;
    return 0;
}

int dns_ensemble_destroy(dns_ensemble* e) {
    delete e;
    return 0;
}

int dns_ensemble_members(const dns_ensemble* e) { return e ? int(e->where.size()) : 0; }

int dns_ensemble_groups(const dns_ensemble* e) { return e ? int(e->e.groups.size()) : 0; }

size_t dns_ensemble_device_bytes(const dns_ensemble* e) { return e ? e->e.bytes() : 0; }

void* dns_ensemble_stream(dns_ensemble* e) { return e ? (void*)e->e.stream() : nullptr; }

static dns_ctx* member_ctx(dns_ensemble* e, int m, int& slot) {
    if (!e || m < 0 || m >= int(e->where.size())) return nullptr;
    slot = e->where[m].second;
    return e->e.groups[e->where[m].first].c;
}

static int ctx_err(dns_ensemble* e, dns_ctx* c, int rc) {
    if (rc) e->e.err = dns_last_error(c);
    return rc;
}

int dns_ensemble_init_fields(dns_ensemble* e) {
    if (!e) return DNS_EINVAL;
    int r = e->e.synchronize();
    if (r) return r;
    for (auto& g : e->e.groups)
        if (g.c && (r = ctx_err(e, g.c, dns_init_fields(g.c)))) return r;
    e->e.steps_done = 0;
    return 0;
}

int dns_ensemble_set_member_fields(dns_ensemble* e, int m, const double* u, const double* v, const double* w,
                                   const double* p) {
    int slot;
    dns_ctx* c = member_ctx(e, m, slot);
    if (!c) return e ? efail(e, DNS_EINVAL, "bad member index") : DNS_EINVAL;
    int r = e->e.synchronize();
    if (r) return r;
    return ctx_err(e, c, dns_set_member_fields(c, slot, u, v, w, p));
}

int dns_ensemble_get_member_fields(dns_ensemble* e, int m, double* u, double* v, double* w, double* p) {
    int slot;
    dns_ctx* c = member_ctx(e, m, slot);
    if (!c) return e ? efail(e, DNS_EINVAL, "bad member index") : DNS_EINVAL;
    int r = e->e.synchronize();
    if (r) return r;
    return ctx_err(e, c, dns_get_member_fields(c, slot, u, v, w, p));
}

int dns_ensemble_member_diagnostics(dns_ensemble* e, int m, dns_diag* d) {
    int slot;
    dns_ctx* c = member_ctx(e, m, slot);
    if (!c) return e ? efail(e, DNS_EINVAL, "bad member index") : DNS_EINVAL;
    int r = e->e.synchronize();
    if (r) return r;
    return ctx_err(e, c, dns_member_diagnostics(c, slot, d));
}

int dns_ensemble_member_shape(const dns_ensemble* e, int m, int* nx, int* ny, int* nz) {
    if (!e || m < 0 || m >= int(e->where.size())) return DNS_EINVAL;
    const dns_ctx* c = e->e.groups[e->where[m].first].c;
    if (nx) *nx = c->nx;
    if (ny) *ny = c->ny;
    if (nz) *nz = c->nz;
    return 0;
}

int dns_ensemble_step(dns_ensemble* e, int nsteps, int use_graph) {
    if (!e) return DNS_EINVAL;
    return e->e.step(nsteps, use_graph != 0);
}

int dns_ensemble_synchronize(dns_ensemble* e) { return e ? e->e.synchronize() : DNS_EINVAL; }

int dns_ensemble_step_timing(dns_ensemble* e, double* out4) {
    if (!e || !out4) return DNS_EINVAL;
    return e->e.timing(out4);
}

int dns_ensemble_set_tuning(dns_ensemble* e, int lanes, int interleave, const int* promote_fp64,
                            const int* promote_fp32) {
    if (!e || lanes < 1 || lanes > ENS_LANES) return DNS_EINVAL;
    int r = e->e.synchronize();
    if (r) return r;
    e->e.tune.lanes = lanes;
    e->e.tune.interleave = interleave ? 1 : 0;
    for (int c = 0; c < FFT_CLASSES; c++) {
        e->e.tune.promote[0][c] = (promote_fp64 && c < 4) ? std::max(0, promote_fp64[c]) : 0;
        e->e.tune.promote[1][c] = (promote_fp32 && c < 4) ? std::max(0, promote_fp32[c]) : 0;
    }
    return e->e.rebuild();
}

int dns_ensemble_set_promotion(dns_ensemble* e, int fp32, const int* levels, int nclasses) {
    static_assert(DNS_ENSEMBLE_SIZE_CLASSES == FFT_CLASSES, "size classes of the ensemble ABI");
    if (!e || (fp32 != 0 && fp32 != 1) || nclasses < 0 || nclasses > FFT_CLASSES || (nclasses > 0 && !levels))
        return DNS_EINVAL;
    int r = e->e.synchronize();
    if (r) return r;
    for (int c = 0; c < FFT_CLASSES; c++) e->e.tune.promote[fp32][c] = c < nclasses ? std::max(0, levels[c]) : 0;
    return e->e.rebuild();
}

int dns_ensemble_launch_counts(dns_ensemble* e, int* counts) {
    if (!e || !counts) return DNS_EINVAL;
    int r = e->e.prepare(false);
    if (r) return r;
    e->e.launch_counts(counts);
    return 0;
}

int dns_ensemble_set_tile_waves(dns_ensemble* e, const int* zwaves) {
    if (!e || !zwaves) return DNS_EINVAL;
    for (int k = 0; k < 3; k++)
        if (zwaves[k] != 0 && zwaves[k] != 2 && zwaves[k] != 4) return efail(e, DNS_EINVAL, "tile waves must be 0, 2 or 4");
    int r = e->e.synchronize();
    if (r) return r;
    for (int k = 0; k < 3; k++) e->e.tune.zwaves[k] = zwaves[k];
    return e->e.rebuild();
}

int dns_ensemble_autotune(dns_ensemble* e, int steps_per_trial, char* report, size_t report_len) {
    if (!e) return DNS_EINVAL;
    std::string rep;
    int r = e->e.autotune(steps_per_trial, rep);
    if (r) return r;
    if (report && report_len > 0) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return 0;
}

int dns_estimate_bytes(const dns_params* p, int nmembers, size_t* bytes) {
    if (!p || nmembers < 1 || !bytes) return DNS_EINVAL;
    std::string err;
    if (!ctx_bytes_for(*p, nmembers, *bytes, err)) { g_ens_error = err; return DNS_ESOLVER; }
    return 0;
}

}  // extern "C"
