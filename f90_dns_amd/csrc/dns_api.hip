// dns_api.hip -- the C ABI of include/f90dns.h: context set-up, resident fields and the
// stage sequence of the 3D fractional step on one MI355X.
//
// State lives in HBM in spectral form: u_hat, v_hat, w_hat, p_hat (complex (nx/2+1, ny, nz)).
// That is exact with respect to the reference, which keeps physical fields and transforms
// them forward at the start of every stage: a real field transformed forward gives a
// Hermitian spectrum, and every operator of the step maps Hermitian spectra to Hermitian
// spectra except the i*k derivatives, whose c2r/r2c round trips the kernels apply as
// effective wavenumbers (x Nyquist column -> 0; y Nyquist row of the kx = 0, nx/2 columns -> 0).
//
// A context holds B >= 1 members of one shape (nx, ny, nz, dt, Lx, Ly, scheme flags); each
// member has its own Re, flow-control targets and initial field.  Members are stored one after
// another in every field array; the plane kernels see them as B*nz planes and the z-kernels
// take the member from blockIdx.z, so one launch per stage advances every member.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/f90dns.h"
#include "../../include/f90dns_debug.h"
#include "context.h"
#include "kernels.h"
#include "setup.h"

using namespace dns;

namespace {
thread_local std::string g_create_error;
}

void dns_set_create_error(const std::string& msg) { g_create_error = msg; }

static int fail(dns_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg; else g_create_error = msg;
    return code;
}
#define HIPCHECK(c, call)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail((c), DNS_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

template <class T>
static int dalloc(dns_ctx* c, T** ptr, size_t n) {
    void* q = nullptr;
    const size_t nb = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, nb);
    if (e != hipSuccess) return fail(c, DNS_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back(q);
    c->bytes += nb;
    *ptr = static_cast<T*>(q);
    return 0;
}
template <class T>
static int upload(dns_ctx* c, T** ptr, const std::vector<T>& h) {
    int r = dalloc(c, ptr, h.size());
    if (r) return r;
    HIPCHECK(c, hipMemcpy(*ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

bool dns_check_gfx950(std::string& why) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { why = "no HIP device"; return false; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) { why = "hipGetDeviceProperties failed"; return false; }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        why = std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950";
        return false;
    }
    return true;
}

bool dns_make_factors(int n, Factors& f) {
    std::vector<int> v;
    if (!radix_factors(n, v) || v.size() > 14) return false;
    f.n = n; f.nf = int(v.size());
    for (size_t i = 0; i < v.size(); i++) f.f[i] = v[i];
    return true;
}

static int destroy_impl(dns_ctx* c) {
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (auto e : c->ev) (void)hipEventDestroy(e);
    for (void* q : c->allocs) (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

static std::string check_params(const dns_params* p) {
    if (p->nx < 8 || p->ny < 4 || (p->nx & 1) || (p->ny & 1)) return "nx must be even and >= 8, ny even and >= 4";
    if (p->nz < 5 || p->nz > 144) return "nz must be in [5, 144]";
    if (!(p->dt > 0) || !(p->re > 0) || !(p->xlen > 0) || !(p->ylen > 0)) return "dt, re, xlen, ylen must be positive";
    if (p->precision != DNS_PREC_FP64 && p->precision != DNS_PREC_MIXED && p->precision != DNS_PREC_MIXED_F32D)
        return "precision must be DNS_PREC_FP64, DNS_PREC_MIXED or DNS_PREC_MIXED_F32D";
    if (p->flow_control_method != 1 && p->flow_control_method != 2) return "flow_control_method must be 1 or 2";
    if (p->flow_control_method == 2 &&
        (p->target_bulk_velocity <= 0 || p->controller_gain <= 0 || p->controller_update_freq <= 0))
        return "method 2 needs positive target_bulk_velocity, controller_gain, controller_update_freq";
    if (p->convection_scheme != 0 && p->convection_scheme != 1) return "convection_scheme must be 0 or 1";
    if (p->pressure_solver != 0 && p->pressure_solver != 1) return "pressure_solver must be 0 or 1";
    if (p->pressure_solver == 1) {
        if (!(p->cgstol >= 0)) return "cgstol must be >= 0";
        if (p->use_weak_form_divergence) return "the jacgs pressure solver is not available with the weak form";
        if (p->precision != DNS_PREC_FP64) return "the jacgs pressure solver runs in FP64 only";
    }
    return "";
}

// members of one context share the grid, the step, the domain and the scheme flags
static bool same_shape(const dns_params& a, const dns_params& b) {
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dt == b.dt && a.xlen == b.xlen && a.ylen == b.ylen &&
           a.use_crank_nicolson == b.use_crank_nicolson && a.use_dealias_23_rule == b.use_dealias_23_rule &&
           a.use_weak_form_divergence == b.use_weak_form_divergence && a.precision == b.precision &&
           a.lgl_derivative == b.lgl_derivative && a.convection_scheme == b.convection_scheme &&
           a.pressure_solver == b.pressure_solver && (a.pressure_solver == 0 || a.cgstol == b.cgstol);
}

extern "C" {

const char* dns_version(void) {
    return "f90_dns_amd 0.2 (gfx950, spectral-resident step, FP64 and mixed precision, batched members)";
}

const char* dns_last_error(const dns_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int dns_create_batch(const dns_params* params, int nmembers, dns_ctx** out) {
    if (!params || !out || nmembers < 1) return fail(nullptr, DNS_EINVAL, "null argument or no members");
    *out = nullptr;
    for (int b = 0; b < nmembers; b++) {
        const std::string e = check_params(&params[b]);
        if (!e.empty()) return fail(nullptr, DNS_EINVAL, nmembers > 1 ? "member " + std::to_string(b) + ": " + e : e);
        if (!same_shape(params[0], params[b]))
            return fail(nullptr, DNS_EINVAL,
                        "member " + std::to_string(b) + " differs from member 0 in grid, dt, domain or scheme flags");
    }
    std::string why;
    if (!dns_check_gfx950(why)) return fail(nullptr, DNS_EDEVICE, why);
    const dns_params* p = &params[0];

    dns_ctx* c = new dns_ctx();
    c->p = *p;
    c->mp.assign(params, params + nmembers);
    c->B = nmembers;
    c->nx = p->nx; c->ny = p->ny; c->nz = p->nz; c->nxh = p->nx / 2 + 1;
    c->H2 = 2 * c->nxh * c->ny; c->npl = c->nx * c->ny;
    c->N = size_t(c->npl) * c->nz; c->NS = size_t(c->H2) * c->nz;
    c->theta = p->use_crank_nicolson ? 0.5 : 1.0;
    c->ybar = 2.0;
    c->rb = (c->nz + 15) / 16;
    c->rbi = (c->nz - 2 + 15) / 16;
    c->mixed = p->precision != DNS_PREC_FP64;
    c->d1f32 = p->precision == DNS_PREC_MIXED_F32D;
    g_create_error.clear();
    auto bail = [&](int code) {
        if (g_create_error.empty()) g_create_error = c->err;
        destroy_impl(c);
        return code;
    };

    if (!fft_geom(c->nx, c->ny, c->g) || (!c->mixed && c->g.cls64 < 0)) {
        g_create_error = "plane too large for the LDS-resident transform: (nx/2)*ny must be <= 8192 in FP64, "
                         "<= 18432 in mixed precision";
        return bail(DNS_EUNSUPPORTED);
    }
    if (!dns_make_factors(c->ny, c->g.fy) || !dns_make_factors(c->nx / 2, c->g.fx)) {
        g_create_error = "nx/2 and ny must be of the form 2^a 3^b";
        return bail(DNS_EINVAL);
    }
    if (fft_prepare(c->g) != hipSuccess) {
        g_create_error = "hipFuncSetAttribute failed for the plane kernels";
        return bail(DNS_EDEVICE);
    }
    if (z_prepare() != hipSuccess) {
        g_create_error = "hipDeviceGetAttribute failed";
        return bail(DNS_EDEVICE);
    }
    // operators: the Helmholtz parts depend on Re and are stacked per member
    std::string err;
    std::vector<double> hF, c1, c2, hl, fv;
    for (int b = 0; b < c->B; b++) {
        Operators ob;
        if (!build_operators(c->nz, p->dt, params[b].re, c->theta, p->use_crank_nicolson != 0,
                             p->lgl_derivative != 0, p->use_weak_form_divergence != 0, ob, err)) {
            g_create_error = err;
            return bail(DNS_ESOLVER);
        }
        if (b == 0) { c->op = ob; c->hf_stride = ob.hF.data.size(); }
        hF.insert(hF.end(), ob.hF.data.begin(), ob.hF.data.end());
        c1.insert(c1.end(), ob.hc1.begin(), ob.hc1.end());
        c2.insert(c2.end(), ob.hc2.begin(), ob.hc2.end());
        hl.insert(hl.end(), ob.hl.begin(), ob.hl.end());
        fv.push_back(ob.f);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        return bail(DNS_EDEVICE);
    }
    int r = 0;
    std::vector<double> tw, xw, yw;
    std::vector<unsigned char> mask;
    twiddles(c->nx, tw);
    std::vector<float> tw32(tw.begin(), tw.end());
    if ((r = upload(c, &c->twx, tw)) || (r = upload(c, &c->twx32, tw32))) return bail(r);
    twiddles(c->ny, tw);
    tw32.assign(tw.begin(), tw.end());
    if ((r = upload(c, &c->twy, tw)) || (r = upload(c, &c->twy32, tw32))) return bail(r);
    c->g.twx = reinterpret_cast<const double2*>(c->twx);
    c->g.twy = reinterpret_cast<const double2*>(c->twy);
    c->g.twx32 = reinterpret_cast<const float2*>(c->twx32);
    c->g.twy32 = reinterpret_cast<const float2*>(c->twy32);
    wavenumbers(c->nx, c->ny, p->xlen, p->ylen, xw, yw);
    dealias_mask(c->nx, c->ny, mask);
    const Operators& o = c->op;
    if ((r = upload(c, &c->xw, xw)) || (r = upload(c, &c->yw, yw)) || (r = upload(c, &c->mask, mask)) ||
        (r = upload(c, &c->wz, o.wz)) || (r = upload(c, &c->e0, o.e0)) || (r = upload(c, &c->f0, o.f0)) ||
        (r = upload(c, &c->D1sw, o.zD1.data)) || (r = upload(c, &c->hF, hF)) || (r = upload(c, &c->hB, o.hB.data)) ||
        (r = upload(c, &c->hc1, c1)) || (r = upload(c, &c->hc2, c2)) || (r = upload(c, &c->hl, hl)) ||
        (r = upload(c, &c->fvec, fv)) || (r = upload(c, &c->pF, o.pF.data)) || (r = upload(c, &c->pB, o.pB.data)) ||
        (r = upload(c, &c->pDB, o.pDB.data)) || (r = upload(c, &c->pl, o.pl)))
        return bail(r);
    if (c->d1f32) {   // the D1 blocks rounded once to FP32 (the FP32 contraction's A operands)
        const std::vector<float> d1f(o.zD1.data.begin(), o.zD1.data.end());
        if ((r = upload(c, &c->D1sw32, d1f))) return bail(r);
    }
    {
        const ZMap z = zmap_make(c->nx, c->ny, ops_wide(c->op));
        std::vector<int4> zt(size_t(2) * z.P);
        std::vector<int> zb(zmap_blocks(z));
        if (!zmap_build(z, reinterpret_cast<int*>(zt.data()), zb.data())) {
            g_create_error = "mode space of the per-mode solves does not come out even";
            return bail(DNS_EINVAL);
        }
        if ((r = upload(c, &c->ztab, zt)) || (r = upload(c, &c->zbfl, zb))) return bail(r);
        c->zm = z;
        c->zm.tab = c->ztab;
        c->zm.bfl = c->zbfl;
    }
    const size_t NSB = c->NS * c->B, NB = c->N * c->B;
    double** specs[] = {&c->uh, &c->vh, &c->wh, &c->ph, &c->us, &c->vs, &c->ws, &c->ox, &c->oy,
                        &c->sh[0], &c->sh[1], &c->sh[2], &c->phi, &c->dphi, &c->tmp};
    for (double** q : specs)
        if ((r = dalloc(c, q, NSB))) return bail(r);
    for (int f = 0; f < 6; f++)
        if ((r = dalloc(c, &c->phys[f], NB))) return bail(r);
    if (c->g.cls64 < 0 && (r = dalloc(c, &c->io, c->NS))) return bail(r);
    if (p->pressure_solver == 1) {
        if ((r = upload(c, &c->stiff, o.stiff)) || (r = dalloc(c, &c->jac_iters, 1)) ||
            (r = dalloc(c, &c->jac_nonconv, 1)))
            return bail(r);
    }
    if ((r = dalloc(c, &c->flow, c->B)) || (r = dalloc(c, &c->fmem, c->B)) ||
        (r = dalloc(c, &c->red_partial, size_t(c->red_blocks) * 5)) || (r = dalloc(c, &c->red_result, 5)))
        return bail(r);
    for (double** q : specs) (void)hipMemset(*q, 0, NSB * sizeof(double));
    std::vector<FlowState> fs(c->B);
    std::vector<FlowMember> fm(c->B);
    c->zwg.assign(c->B, 0);
    for (int b = 0; b < c->B; b++) {
        const dns_params& q = params[b];
        fs[b] = FlowState{};
        fs[b].G = q.target_pressure_gradient;
        fs[b].gavg = q.target_pressure_gradient;
        fm[b] = FlowMember{q.target_pressure_gradient, q.target_bulk_velocity, q.controller_gain, q.re,
                           q.flow_control_method, q.controller_update_freq, q.istart, 0};
    }
    if (hipMemcpy(c->flow, fs.data(), sizeof(FlowState) * c->B, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->fmem, fm.data(), sizeof(FlowMember) * c->B, hipMemcpyHostToDevice) != hipSuccess) {
        g_create_error = "upload of the flow-control state failed";
        return bail(DNS_EDEVICE);
    }
    *out = c;
    return 0;
}

int dns_create(const dns_params* p, dns_ctx** out) { return dns_create_batch(p, 1, out); }

int dns_destroy(dns_ctx* c) { return destroy_impl(c); }

void* dns_stream(dns_ctx* c) { return c ? (void*)c->stream : nullptr; }

int dns_members(const dns_ctx* c) { return c ? c->B : 0; }

size_t dns_device_bytes(const dns_ctx* c) { return c ? c->bytes : 0; }

static FoldOp fold_op(const FoldedA& f, const double* dev) {
    FoldOp o;
    o.A = dev; o.Af = nullptr; o.rb0 = f.rb0; o.rb1 = f.rb1; o.ks = f.ks; o.ksp = f.ksp; o.rows0 = f.rows0; o.rows1 = f.rows1;
    o.extra = f.extra;
    o.pb = f.pb;
    return o;
}

extern "C++" bool dns::ops_wide(const Operators& o) {
    const FoldOp zd = fold_op(o.zD1, nullptr), hf = fold_op(o.hF, nullptr), hb = fold_op(o.hB, nullptr);
    const FoldOp pf = fold_op(o.pF, nullptr), pb = fold_op(o.pB, nullptr), pdb = fold_op(o.pDB, nullptr);
    const FoldOp* z1[1] = {&zd};
    const FoldOp* h2[2] = {&hf, &hb};
    const FoldOp* p3[3] = {&pf, &pb, &pdb};
    const int v[3] = {z_variant_of(z1, 1), z_variant_of(h2, 2), z_variant_of(p3, 3)};
    for (int x : v)
        if (x < 3) return false;
    return true;
}

static FFTArgs fft_args(dns_ctx* c) {
    FFTArgs a;
    std::memset(&a, 0, sizeof(a));
    a.xw = c->xw; a.yw = c->yw;
    return a;
}

// transforms of nplanes <= nz consecutive planes (device pointers, FP64 storage, FP64
// arithmetic; two passes through c->io where the plane does not fit the LDS in FP64)
static int r2c_planes(dns_ctx* c, const double* src, double* dst, int nplanes) {
    if (c->g.cls64 < 0) {
        HIPCHECK(c, launch_two_pass(c->g, true, src, dst, c->io, nplanes, c->npl, c->H2, c->stream));
        return 0;
    }
    FFTArgs a = fft_args(c);
    a.src[0] = src; a.dst[0] = dst;
    a.sstride = c->npl; a.dstride = c->H2;
    HIPCHECK(c, launch_r2c(c->g, R2C_PLAIN, a, nplanes, 1, false, c->stream));
    return 0;
}
static int c2r_planes(dns_ctx* c, const double* src, double* dst, int nplanes) {
    if (c->g.cls64 < 0) {
        HIPCHECK(c, launch_two_pass(c->g, false, src, dst, c->io, nplanes, c->H2, c->npl, c->stream));
        return 0;
    }
    FFTArgs a = fft_args(c);
    a.src[0] = src; a.dst[0] = dst;
    a.sstride = c->H2; a.dstride = c->npl;
    HIPCHECK(c, launch_c2r(c->g, C2R_PLAIN, a, nplanes, 1, false, c->stream));
    return 0;
}

static bool bad_member(const dns_ctx* c, int b) { return !c || b < 0 || b >= c->B; }

// A pressure field set from outside (a restart read) leaves the wall pressure gradient of the
// Helmholtz boundary values at zero until the next pressure correction, as grad_p_prev is in
// the reference (zeroed by initialize_fields_3d :957 and written only by pressure_correction_3d
// :1923; read_restart_file_3d does not restore it).
static int mark_zero_wallgrad(dns_ctx* c, int b) {
    const int one = 1;
    HIPCHECK(c, hipMemcpyAsync(reinterpret_cast<char*>(c->flow + b) + offsetof(FlowState, zero_wallgrad), &one,
                               sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->zwg[b] = 1;
    return 0;
}

// after a pressure correction on the stream: the wall gradients follow p again
extern "C++" int dns::ctx_after_correction(dns_ctx* c, hipStream_t s) {
    bool any = false;
    for (char f : c->zwg) any = any || f;
    if (!any) return 0;
    HIPCHECK(c, launch_clear_wallgrad(c->flow, c->B, s));
    c->zwg.assign(c->B, 0);
    return 0;
}

int dns_set_member_fields(dns_ctx* c, int b, const double* u, const double* v, const double* w, const double* p) {
    if (bad_member(c, b)) return DNS_EINVAL;
    const double* in[4] = {u, v, w, p};
    double* outs[4] = {c->uh, c->vh, c->wh, c->ph};
    for (int f = 0; f < 4; f++) {
        if (!in[f]) continue;
        HIPCHECK(c, hipMemcpyAsync(c->phys[0], in[f], c->N * sizeof(double), hipMemcpyHostToDevice, c->stream));
        int r = r2c_planes(c, c->phys[0], outs[f] + size_t(b) * c->NS, c->nz);
        if (r) return r;
    }
    if (p) {
        int r = mark_zero_wallgrad(c, b);
        if (r) return r;
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_get_member_fields(dns_ctx* c, int b, double* u, double* v, double* w, double* p) {
    if (bad_member(c, b)) return DNS_EINVAL;
    double* outs[4] = {u, v, w, p};
    const double* in[4] = {c->uh, c->vh, c->wh, c->ph};
    for (int f = 0; f < 4; f++) {
        if (!outs[f]) continue;
        int r = c2r_planes(c, in[f] + size_t(b) * c->NS, c->phys[f], c->nz);
        if (r) return r;
        HIPCHECK(c, hipMemcpyAsync(outs[f], c->phys[f], c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_set_fields(dns_ctx* c, const double* u, const double* v, const double* w, const double* p) {
    return dns_set_member_fields(c, 0, u, v, w, p);
}

int dns_get_fields(dns_ctx* c, double* u, double* v, double* w, double* p) {
    return dns_get_member_fields(c, 0, u, v, w, p);
}

int dns_set_fields_device(dns_ctx* c, const double* u, const double* v, const double* w, const double* p) {
    if (!c) return DNS_EINVAL;
    const double* in[4] = {u, v, w, p};
    double* outs[4] = {c->uh, c->vh, c->wh, c->ph};
    for (int f = 0; f < 4; f++) {
        if (!in[f]) continue;
        int r = r2c_planes(c, in[f], outs[f], c->nz);
        if (r) return r;
    }
    if (p) {
        int r = mark_zero_wallgrad(c, 0);
        if (r) return r;
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_get_fields_device(dns_ctx* c, double* u, double* v, double* w, double* p) {
    if (!c) return DNS_EINVAL;
    double* outs[4] = {u, v, w, p};
    const double* in[4] = {c->uh, c->vh, c->wh, c->ph};
    for (int f = 0; f < 4; f++) {
        if (!outs[f]) continue;
        int r = c2r_planes(c, in[f], outs[f], c->nz);
        if (r) return r;
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_init_member_fields(dns_ctx* c, int b);

int dns_init_fields(dns_ctx* c) {
    if (!c) return DNS_EINVAL;
    for (int b = 0; b < c->B; b++) {
        int r = dns_init_member_fields(c, b);
        if (r) return r;
    }
    c->time = 0.0; c->step = 0;
    return 0;
}

int dns_set_member_controller(dns_ctx* c, int b, double G, double integral, double prev, double gsum, int gcount) {
    if (bad_member(c, b)) return DNS_EINVAL;
    FlowState fs{};
    fs.G = G; fs.err_int = integral; fs.err_prev = prev; fs.gsum = gsum; fs.gcount = gcount;
    fs.gavg = gcount > 0 ? gsum / double(gcount) : G;
    fs.zero_wallgrad = c->zwg[b];
    HIPCHECK(c, hipMemcpyAsync(c->flow + b, &fs, sizeof(fs), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_set_member_params(dns_ctx* c, int b, const dns_params* q) {
    if (bad_member(c, b) || !q) return DNS_EINVAL;
    const std::string e = check_params(q);
    if (!e.empty()) return fail(c, DNS_EINVAL, e);
    if (!same_shape(c->p, *q)) return fail(c, DNS_EINVAL, "member parameters differ in grid, dt, domain or scheme flags");
    Operators ob;
    std::string err;
    if (!build_operators(c->nz, c->p.dt, q->re, c->theta, c->p.use_crank_nicolson != 0, c->p.lgl_derivative != 0,
                         c->p.use_weak_form_divergence != 0, ob, err))
        return fail(c, DNS_ESOLVER, err);
    const int m = c->nz - 2;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(c->hF + size_t(b) * c->hf_stride, ob.hF.data.data(), c->hf_stride * sizeof(double),
                          hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->hc1 + size_t(b) * m, ob.hc1.data(), m * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->hc2 + size_t(b) * m, ob.hc2.data(), m * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->hl + size_t(b) * m, ob.hl.data(), m * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->fvec + b, &ob.f, sizeof(double), hipMemcpyHostToDevice));
    FlowState fs{};
    fs.G = q->target_pressure_gradient;
    fs.gavg = q->target_pressure_gradient;
    c->zwg[b] = 0;
    FlowMember fm{q->target_pressure_gradient, q->target_bulk_velocity, q->controller_gain, q->re,
                  q->flow_control_method, q->controller_update_freq, q->istart, 0};
    HIPCHECK(c, hipMemcpy(c->flow + b, &fs, sizeof(fs), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->fmem + b, &fm, sizeof(fm), hipMemcpyHostToDevice));
    c->mp[b] = *q;
    if (b == 0) c->p = *q;
    return 0;
}

// initialize_fields_3d for one member: the separable form is built on the host (planar
// field, wall-normal profile, amplitude scale), the 3D fields on the device
int dns_init_member_fields(dns_ctx* c, int b) {
    if (bad_member(c, b)) return DNS_EINVAL;
    const dns_params& q = c->mp[b];
    InitSpec s{c->nx, c->ny, c->nz, q.xlen, q.ylen, q.u_centerline, q.enable_perturbations,
               q.perturbation_kind, q.analytical_mode_x, q.analytical_mode_y, q.perturbation_amplitude,
               (uint64_t)q.seed, q.noise_kmax};
    PlanarInit pl;
    initial_planar(s, pl);
    // staging in the physical work buffer 3 (B*N >= 2*npl + 2*nz doubles)
    double* st = c->phys[3];
    const size_t npl = size_t(c->npl);
    HIPCHECK(c, hipMemcpyAsync(st, pl.u2.data(), npl * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(st + npl, pl.v2.data(), npl * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(st + 2 * npl, pl.U.data(), c->nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(st + 2 * npl + c->nz, pl.wd.data(), c->nz * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
    InitArgs a{};
    a.npl = c->npl; a.nz = c->nz;
    a.u2 = st; a.v2 = st + npl; a.U = st + 2 * npl; a.wd = st + 2 * npl + c->nz;
    a.sf = pl.sf; a.scaled = pl.scaled; a.perturbed = pl.perturbed;
    a.u = c->phys[0]; a.v = c->phys[1]; a.w = c->phys[2];
    HIPCHECK(c, launch_init_fields(a, c->stream));
    const size_t mo = size_t(b) * c->NS;
    int r;
    if ((r = r2c_planes(c, c->phys[0], c->uh + mo, c->nz)) || (r = r2c_planes(c, c->phys[1], c->vh + mo, c->nz)) ||
        (r = r2c_planes(c, c->phys[2], c->wh + mo, c->nz)))
        return r;
    HIPCHECK(c, hipMemsetAsync(c->ph + mo, 0, c->NS * sizeof(double), c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_set_controller(dns_ctx* c, double G, double integral, double prev, double gsum, int gcount) {
    return dns_set_member_controller(c, 0, G, integral, prev, gsum, gcount);
}

// ---- stages -----------------------------------------------------------------------------
int dns_begin_step(dns_ctx* c) {
    if (!c) return DNS_EINVAL;
    c->step += 1;
    c->time = c->time + c->p.dt;
    return 0;
}

extern "C++" FlowArgs dns::ctx_flow_args(dns_ctx* c) {
    FlowArgs fa{};
    fa.nz = c->nz; fa.dt = c->p.dt; fa.nxny = double(c->nx * c->ny);
    double ws = 0.0;
    for (int k = 0; k < c->nz; k++) ws += c->op.wz[k];
    fa.wsum = ws;
    fa.uh = c->uh; fa.H2 = c->H2; fa.mstride = (long long)c->NS; fa.wz = c->wz;
    fa.mem = c->fmem; fa.st = c->flow;
    return fa;
}

extern "C++" ZArgs dns::ctx_vort_args(dns_ctx* c) {
    ZArgs z{};
    z.mstride = (long long)c->NS;
    z.nz = c->nz; z.H2 = c->H2; z.nxh = c->nxh; z.ny = c->ny; z.nx = c->nx;
    z.d1 = fold_op(c->op.zD1, c->D1sw);
    z.in[0] = c->uh; z.in[1] = c->vh; z.in[2] = c->wh;
    z.out[0] = c->ox; z.out[1] = c->oy;
    z.xw = c->xw; z.yw = c->yw;
    z.out_f32 = c->mixed ? 1 : 0;
    z.d1.Af = c->D1sw32;
    z.d1_f32 = c->d1f32 ? 1 : 0;
    z.packed = 1;
    z.zm = c->zm;
    return z;
}

// the nonlinear term's inputs: the state (u, v, w), or a convection stage's velocities
struct NLInputs { const double* f[3]; };
static NLInputs nl_inputs(const dns_ctx* c, bool stage) {
    return stage ? NLInputs{{c->us, c->vs, c->ws}} : NLInputs{{c->uh, c->vh, c->wh}};
}

static int nl_zderiv(dns_ctx* c, bool stage = false) {
    ZArgs z = ctx_vort_args(c);
    const NLInputs in = nl_inputs(c, stage);
    for (int f = 0; f < 3; f++) z.in[f] = in.f[f];
    HIPCHECK(c, launch_zderiv(Z_VORT, z, 1, c->B, c->stream));
    return 0;
}

extern "C++" FFTArgs dns::ctx_c2r_nl_args(dns_ctx* c) {
    FFTArgs a = fft_args(c);
    a.src[0] = c->uh; a.src[1] = c->vh; a.src[2] = c->wh; a.src[3] = c->ox; a.src[4] = c->oy;
    for (int f = 0; f < 6; f++) a.dst[f] = c->phys[f];
    a.sstride = c->H2; a.dstride = c->npl;
    a.src_f32 = c->mixed ? 0x18u : 0u;
    a.dst_f32 = c->mixed ? 1 : 0;
    return a;
}

// The fused step transforms the interior planes only: the update uses S = u x omega at the
// interior nodes alone (the wall rows of the Helmholtz solve are the boundary values), so the
// wall planes of the six physical fields and of S are not needed.  The stage API
// (dns_compute_source_terms) computes every plane, as the reference does.
static int nl_c2r(dns_ctx* c, bool walls, bool stage = false) {
    FFTArgs a = ctx_c2r_nl_args(c);
    const NLInputs in = nl_inputs(c, stage);
    for (int f = 0; f < 3; f++) a.src[f] = in.f[f];
    a.zin = walls ? 0 : c->nz - 2;
    HIPCHECK(c, launch_c2r(c->g, C2R_NL, a, (walls ? c->nz : c->nz - 2) * c->B, 6, c->mixed, c->stream));
    return 0;
}

extern "C++" FFTArgs dns::ctx_r2c_cross_args(dns_ctx* c) {
    FFTArgs b = fft_args(c);
    for (int f = 0; f < 6; f++) b.src[f] = c->phys[f];
    for (int f = 0; f < 3; f++) b.dst[f] = c->sh[f];
    b.mask = c->p.use_dealias_23_rule ? c->mask : nullptr;
    b.sstride = c->npl; b.dstride = c->H2;
    b.src_f32 = c->mixed ? 0x3Fu : 0u;
    b.dst_f32 = c->mixed ? 1 : 0;
    return b;
}

static int nl_r2c(dns_ctx* c, bool walls) {
    FFTArgs a = ctx_r2c_cross_args(c);
    a.zin = walls ? 0 : c->nz - 2;
    HIPCHECK(c, launch_r2c(c->g, R2C_CROSS, a, (walls ? c->nz : c->nz - 2) * c->B, 3, c->mixed, c->stream));
    return 0;
}

static int flow_control(dns_ctx* c) {
    HIPCHECK(c, launch_flow_control(ctx_flow_args(c), c->B, c->stream));
    return 0;
}

// The 4-stage 1/irk convection of runge_kutta_convection (DNS_pressure_BC_2D.f90:931-1014) on
// the spectral state: stage irk = 4 takes S of (u, v, w); each later stage takes S of
// un + (dt/(irk+1)) S_prev, held in the star buffers (written by the Helmholtz launch only
// afterwards), walls zero; the S of the last stage stays in the S buffers.  The state itself is
// not touched, which is the reference's reset to un.
static int rk_convection(dns_ctx* c, bool walls) {
    int r;
    for (int irk = 4; irk >= 1; irk--) {
        const bool stage = irk < 4;
        if ((r = nl_zderiv(c, stage)) || (r = nl_c2r(c, walls, stage)) || (r = nl_r2c(c, walls))) return r;
        if (irk == 1) break;
        RKStageArgs a{};
        a.n = (long long)c->NS * c->B; a.H2 = c->H2; a.nz = c->nz;
        a.un[0] = c->uh; a.un[1] = c->vh; a.un[2] = c->wh;
        for (int f = 0; f < 3; f++) a.s[f] = c->sh[f];
        a.s_f32 = c->mixed ? 1 : 0;
        a.out[0] = c->us; a.out[1] = c->vs; a.out[2] = c->ws;
        a.a = (1.0 / double(irk)) * c->p.dt;   // alf * dt of this stage (:985-991)
        HIPCHECK(c, launch_rk_stage(a, c->stream));
    }
    return 0;
}

// flow: launch the flow controller here (the stage API); the fused step runs it inside the
// Helmholtz launch instead
static int source_terms(dns_ctx* c, bool walls, bool flow) {
    int r;
    if (c->p.convection_scheme == 1) {
        if ((r = rk_convection(c, walls))) return r;
    } else if ((r = nl_zderiv(c)) || (r = nl_c2r(c, walls)) || (r = nl_r2c(c, walls))) {
        return r;
    }
    return flow ? flow_control(c) : 0;
}

int dns_compute_source_terms(dns_ctx* c) {
    if (!c) return DNS_EINVAL;
    return source_terms(c, true, true);
}

extern "C++" HelmArgs dns::ctx_helm_args(dns_ctx* c) {
    HelmArgs h{};
    h.mstride = (long long)c->NS; h.fstride = (long long)c->hf_stride;
    h.nz = c->nz; h.m = c->nz - 2; h.H2 = c->H2; h.nxh = c->nxh; h.ny = c->ny; h.nx = c->nx;
    h.hF = fold_op(c->op.hF, c->hF); h.hB = fold_op(c->op.hB, c->hB);
    h.c1 = c->hc1; h.c2 = c->hc2; h.hl = c->hl; h.fvec = c->fvec;
    h.dt = c->p.dt; h.nxny = double(c->nx * c->ny);
    h.cn = c->p.use_crank_nicolson ? 1 : 0;
    h.xw = c->xw; h.yw = c->yw;
    h.uh[0] = c->uh; h.uh[1] = c->vh; h.uh[2] = c->wh;
    for (int f = 0; f < 3; f++) h.sh[f] = c->sh[f];
    h.out[0] = c->us; h.out[1] = c->vs; h.out[2] = c->ws;
    h.ph = c->ph;
    h.s_f32 = c->mixed ? 1 : 0;
    h.flow = c->flow;
    h.zm = c->zm;
    return h;
}

// with_flow: the Helmholtz launch also runs the flow controller (the fused step)
static int viscous(dns_ctx* c, bool with_flow) {
    HelmArgs h = ctx_helm_args(c);
    if (with_flow) {
        h.fl = ctx_flow_args(c);
        h.flow_fused = 1;
    }
    HIPCHECK(c, launch_helmholtz(h, c->B, c->stream));
    return 0;
}

int dns_viscous_step(dns_ctx* c) {
    if (!c) return DNS_EINVAL;
    return viscous(c, false);
}

extern "C++" PoisArgs dns::ctx_pois_args(dns_ctx* c, int fused) {
    PoisArgs a{};
    a.mstride = (long long)c->NS;
    a.nz = c->nz; a.H2 = c->H2; a.nxh = c->nxh; a.ny = c->ny; a.nx = c->nx;
    a.pF = fold_op(c->op.pF, c->pF); a.pB = fold_op(c->op.pB, c->pB); a.pDB = fold_op(c->op.pDB, c->pDB);
    a.pl = c->pl; a.dt = c->p.dt; a.fused = fused; a.weak = c->p.use_weak_form_divergence != 0; a.xw = c->xw; a.yw = c->yw;
    a.us = c->us; a.vs = c->vs; a.ws = c->ws;
    a.u = c->uh; a.v = c->vh; a.w = c->wh; a.p = c->ph;
    a.phi = c->phi; a.dphi = c->dphi;
    a.m0 = ctx_mode0_args(c, fused);   // mode (0,0) solved by the Poisson launch
    a.mode0 = 1;
    a.zm = c->zm;
    return a;
}

extern "C++" Mode0Args dns::ctx_mode0_args(dns_ctx* c, int fused) {
    Mode0Args m{};
    m.mstride = (long long)c->NS;
    m.nz = c->nz; m.H2 = c->H2; m.e0 = c->e0; m.f0 = c->f0; m.dt = c->p.dt; m.fused = fused;
    m.us = c->us; m.vs = c->vs; m.ws = c->ws;
    m.u = c->uh; m.v = c->vh; m.w = c->wh; m.p = c->ph;
    m.phi = c->phi; m.dphi = c->dphi;
    return m;
}

// plain D1 contraction of nf = 1 spectral field over all members: out = D1 in
static int zderiv_plain(dns_ctx* c, const double* in, double* out) {
    ZArgs z{};
    z.mstride = (long long)c->NS;
    z.nz = c->nz; z.H2 = c->H2; z.nxh = c->nxh; z.ny = c->ny; z.nx = c->nx;
    z.d1 = fold_op(c->op.zD1, c->D1sw);
    z.in[0] = in; z.out[0] = out; z.xw = c->xw; z.yw = c->yw;
    HIPCHECK(c, launch_zderiv(Z_PLAIN, z, 1, c->B, c->stream));
    return 0;
}

static int correction(dns_ctx* c) {
    CorrArgs a{};
    a.nz = c->nz * c->B;   // members are consecutive planes and the update is per element
    a.H2 = c->H2; a.nxh = c->nxh; a.ny = c->ny; a.nx = c->nx;
    a.xw = c->xw; a.yw = c->yw;
    a.us = c->us; a.vs = c->vs; a.ws = c->ws; a.phi = c->phi; a.dphi = c->dphi;
    a.u = c->uh; a.v = c->vh; a.w = c->wh; a.p = c->ph; a.dt = c->p.dt;
    HIPCHECK(c, launch_correction(a, c->stream));
    return 0;
}

// pressure_solver 1: phi from jacgs per real column (RHS from u*, v* and D1 w*), D1 phi; when
// fused also the projection, p += phi and the walls
static int pressure_jacgs(dns_ctx* c, int fused) {
    int r;
    if ((r = zderiv_plain(c, c->ws, c->tmp))) return r;
    HIPCHECK(c, launch_jacgs_clear(c->jac_iters, c->jac_nonconv, c->stream));
    JacgsArgs j{};
    j.mstride = (long long)c->NS;
    j.nz = c->nz; j.H2 = c->H2; j.nxh = c->nxh; j.nx = c->nx; j.ny = c->ny; j.niter = 500;
    j.st = c->stiff; j.wz = c->wz; j.xw = c->xw; j.yw = c->yw;
    j.us = c->us; j.vs = c->vs; j.dws = c->tmp; j.phi = c->phi;
    j.dt = c->p.dt; j.ybar = c->ybar; j.tol = c->p.cgstol;
    j.iters = c->jac_iters; j.nonconv = c->jac_nonconv;
    HIPCHECK(c, launch_jacgs(j, c->B, c->stream));
    if ((r = zderiv_plain(c, c->phi, c->dphi))) return r;
    if (!fused) return 0;
    if ((r = correction(c))) return r;
    HIPCHECK(c, launch_zero_walls(c->nz, c->H2, (long long)c->NS, c->B, c->uh, c->vh, c->wh, c->stream));
    return 0;
}

static int pressure(dns_ctx* c, int fused) {
    if (c->p.pressure_solver == 1) return pressure_jacgs(c, fused);
    PoisArgs a = ctx_pois_args(c, fused);
    HIPCHECK(c, launch_poisson(a, c->B, c->stream));
    return 0;
}

int dns_solve_pressure(dns_ctx* c) {
    if (!c) return DNS_EINVAL;
    return pressure(c, 0);
}

int dns_pressure_correction(dns_ctx* c) {
    if (!c) return DNS_EINVAL;
    int r = correction(c);
    return r ? r : ctx_after_correction(c, c->stream);
}

int dns_apply_boundary_conditions(dns_ctx* c) {
    if (!c) return DNS_EINVAL;
    HIPCHECK(c, launch_zero_walls(c->nz, c->H2, (long long)c->NS, c->B, c->uh, c->vh, c->wh, c->stream));
    return 0;
}

// the captured step: no host logic after the launches
static int one_step_kernels(dns_ctx* c) {
    int r;
    if ((r = source_terms(c, false, false))) return r;
    if ((r = viscous(c, true))) return r;
    return pressure(c, 1);
}
static int one_step_fused(dns_ctx* c) {
    int r = one_step_kernels(c);
    return r ? r : ctx_after_correction(c, c->stream);
}

static int ensure_events(dns_ctx* c, int n) {
    while ((int)c->ev.size() < n) {
        hipEvent_t e;
        HIPCHECK(c, hipEventCreate(&e));
        c->ev.push_back(e);
    }
    return 0;
}

static int collect_timing(dns_ctx* c, int nsteps) {
    HIPCHECK(c, hipEventSynchronize(c->ev[nsteps]));
    c->step_ms.resize(nsteps);
    for (int s = 0; s < nsteps; s++) {
        float ms = 0.f;
        HIPCHECK(c, hipEventElapsedTime(&ms, c->ev[s], c->ev[s + 1]));
        c->step_ms[s] = ms;
    }
    return 0;
}

int dns_step(dns_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return DNS_EINVAL;
    int r = ensure_events(c, nsteps + 1);
    if (r) return r;
    HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    for (int s = 0; s < nsteps; s++) {
        dns_begin_step(c);
        if ((r = one_step_fused(c))) return r;
        HIPCHECK(c, hipEventRecord(c->ev[s + 1], c->stream));
    }
    c->step_ms.clear();
    c->step_ms.push_back(-double(nsteps));   // marker: timing not yet collected
    return 0;
}

// capture one fused step without executing it
static int ensure_graph(dns_ctx* c) {
    if (c->gexec) return 0;
    HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int r = one_step_kernels(c);
    hipGraph_t gph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &gph);
    if (r) {
        if (gph) (void)hipGraphDestroy(gph);
        return r;
    }
    if (e != hipSuccess) return fail(c, DNS_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    c->graph = gph;
    HIPCHECK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    return 0;
}

int dns_step_graph(dns_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return DNS_EINVAL;
    int r;
    if ((r = ensure_graph(c))) return r;
    if ((r = ensure_events(c, nsteps + 1))) return r;
    HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    for (int s = 0; s < nsteps; s++) {
        dns_begin_step(c);
        HIPCHECK(c, hipGraphLaunch(c->gexec, c->stream));
        if ((r = ctx_after_correction(c, c->stream))) return r;
        HIPCHECK(c, hipEventRecord(c->ev[s + 1], c->stream));
    }
    c->step_ms.clear();
    c->step_ms.push_back(-double(nsteps));
    return 0;
}

int dns_enqueue_step(dns_ctx* c, int use_graph) {
    if (!c) return DNS_EINVAL;
    int r;
    if (use_graph && (r = ensure_graph(c))) return r;
    dns_begin_step(c);
    if (use_graph) {
        HIPCHECK(c, hipGraphLaunch(c->gexec, c->stream));
        return ctx_after_correction(c, c->stream);
    }
    return one_step_fused(c);
}

int dns_synchronize(dns_ctx* c) {
    if (!c) return DNS_EINVAL;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_query(dns_ctx* c) {
    if (!c) return DNS_EINVAL;
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    return fail(c, DNS_EDEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int dns_step_timing(dns_ctx* c, double* out4) {
    if (!c || !out4) return DNS_EINVAL;
    if (!c->step_ms.empty() && c->step_ms[0] < 0) {
        const int n = int(-c->step_ms[0]);
        if (n == 0) c->step_ms.clear();
        else {
            int r = collect_timing(c, n);
            if (r) return r;
        }
    }
    if (c->step_ms.empty()) { out4[0] = out4[1] = out4[2] = out4[3] = 0.0; return 0; }
    std::vector<double> t = c->step_ms;
    std::sort(t.begin(), t.end());
    auto pct = [&](double q) {
        const double pos = q * (t.size() - 1);
        const size_t lo = size_t(std::floor(pos)), hi = std::min(t.size() - 1, lo + 1);
        return t[lo] + (pos - lo) * (t[hi] - t[lo]);
    };
    double sum = 0.0;
    for (double x : t) sum += x;
    out4[0] = t.front(); out4[1] = pct(0.5); out4[2] = pct(0.99); out4[3] = sum / t.size();
    return 0;
}

// ---- diagnostics --------------------------------------------------------------------------
static int reduce(dns_ctx* c, const double* u, const double* v, const double* w, const double* d, double* res) {
    ReduceArgs a{};
    a.npl = c->npl; a.nz = c->nz; a.u = u; a.v = v; a.w = w; a.d = d; a.wz = c->wz;
    a.partial = c->red_partial; a.result = c->red_result; a.nblocks = c->red_blocks;
    HIPCHECK(c, launch_reduce(a, c->stream));
    HIPCHECK(c, hipMemcpyAsync(res, c->red_result, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_member_diagnostics(dns_ctx* c, int b, dns_diag* d) {
    if (bad_member(c, b) || !d) return DNS_EINVAL;
    const size_t mo = size_t(b) * c->NS;
    const double* in[3] = {c->uh + mo, c->vh + mo, c->wh + mo};
    for (int f = 0; f < 3; f++) {
        int r = c2r_planes(c, in[f], c->phys[f], c->nz);
        if (r) return r;
    }
    double res[5];
    int r = reduce(c, c->phys[0], c->phys[1], c->phys[2], nullptr, res);
    if (r) return r;
    std::vector<double> u00(c->nz);
    HIPCHECK(c, hipMemcpy2DAsync(u00.data(), sizeof(double), c->uh + mo, size_t(c->H2) * sizeof(double),
                                 sizeof(double), c->nz, hipMemcpyDeviceToHost, c->stream));
    FlowState fs;
    HIPCHECK(c, hipMemcpyAsync(&fs, c->flow + b, sizeof(fs), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    const double xlen = c->p.xlen, ylen = c->p.ylen, ybar = c->ybar;
    const double fac = (xlen / double(c->nx)) * (ylen / double(c->ny)) * (ybar / 2.0);
    d->umax = res[0]; d->vmax = res[1]; d->wmax = res[2];
    d->energy = res[3] * fac / (xlen * ylen * ybar);
    double vol = 0.0, ws = 0.0;
    for (int k = 0; k < c->nz; k++) { vol += u00[k] * c->op.wz[k]; ws += c->op.wz[k]; }
    d->bulk_velocity = vol / (double(c->nx * c->ny) * ws);
    d->pressure_gradient = fs.G;
    d->avg_pressure_gradient = fs.gavg;
    d->bulk_error_integral = fs.err_int;
    d->bulk_error_prev = fs.err_prev;
    d->time = c->time;
    d->step = c->step;
    return 0;
}

int dns_diagnostics(dns_ctx* c, dns_diag* d) { return dns_member_diagnostics(c, 0, d); }

int dns_member_divergence(dns_ctx* c, int b, double* max_div) {
    if (bad_member(c, b) || !max_div) return DNS_EINVAL;
    const size_t mo = size_t(b) * c->NS;
    ZArgs z{};
    z.mstride = 0;
    z.nz = c->nz; z.H2 = c->H2; z.nxh = c->nxh; z.ny = c->ny; z.nx = c->nx;
    z.d1 = fold_op(c->op.zD1, c->D1sw);
    z.in[0] = c->wh + mo; z.out[0] = c->tmp; z.xw = c->xw; z.yw = c->yw;
    HIPCHECK(c, launch_zderiv(Z_PLAIN, z, 1, 1, c->stream));
    if (c->g.cls64 < 0) {
        // combine into the omega_x buffer (free outside dns_compute_source_terms), then two passes
        HIPCHECK(c, launch_div_spec(c->g, c->xw, c->yw, c->uh + mo, c->vh + mo, c->tmp, c->ox, c->nz, c->stream));
        int r = c2r_planes(c, c->ox, c->phys[0], c->nz);
        if (r) return r;
    } else {
        FFTArgs a = fft_args(c);
        a.src[0] = c->uh + mo; a.src[1] = c->vh + mo; a.src[2] = c->tmp; a.dst[0] = c->phys[0];
        a.sstride = c->H2; a.dstride = c->npl;
        HIPCHECK(c, launch_c2r(c->g, C2R_DIV, a, c->nz, 1, false, c->stream));
    }
    double res[5];
    int r = reduce(c, nullptr, nullptr, nullptr, c->phys[0], res);
    if (r) return r;
    *max_div = res[4];
    return 0;
}

int dns_check_divergence(dns_ctx* c, double* max_div) { return dns_member_divergence(c, 0, max_div); }

int dns_grid_nodes(const dns_ctx* c, double* z, double* w) {
    if (!c) return DNS_EINVAL;
    if (z) std::copy(c->op.z.begin(), c->op.z.end(), z);
    if (w) std::copy(c->op.wz.begin(), c->op.wz.end(), w);
    return 0;
}

// ---- debug hooks (include/f90dns_debug.h) -------------------------------------------------
int dns_debug_spectral(dns_ctx* c, int which, double* out) {
    if (!c || !out) return DNS_EINVAL;
    const double* src[] = {c->uh, c->vh, c->wh, c->ph, c->us, c->vs, c->ws, c->sh[0], c->sh[1], c->sh[2],
                           c->ox, c->oy, c->phi, c->dphi};
    if (which < 0 || which >= int(sizeof(src) / sizeof(src[0]))) return fail(c, DNS_EINVAL, "bad buffer index");
    HIPCHECK(c, hipMemcpyAsync(out, src[which], c->NS * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_debug_zderiv(dns_ctx* c, const double* spec_in, double* spec_out) {
    if (!c || !spec_in || !spec_out) return DNS_EINVAL;
    HIPCHECK(c, hipMemcpyAsync(c->tmp, spec_in, c->NS * sizeof(double), hipMemcpyHostToDevice, c->stream));
    ZArgs z{};
    z.mstride = 0;
    z.nz = c->nz; z.H2 = c->H2; z.nxh = c->nxh; z.ny = c->ny; z.nx = c->nx;
    z.d1 = fold_op(c->op.zD1, c->D1sw);
    z.in[0] = c->tmp; z.out[0] = c->phi; z.xw = c->xw; z.yw = c->yw;
    HIPCHECK(c, launch_zderiv(Z_PLAIN, z, 1, 1, c->stream));
    HIPCHECK(c, hipMemcpyAsync(spec_out, c->phi, c->NS * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// One eager fused step per iteration with an event between every kernel launch; out[i] is
// the mean duration (ms) of kernel i in the order of dns_profile_kernel_name.
static const char* kKernelNames[] = {"zderiv_vort", "c2r_nl", "r2c_cross", "flow_control", "helmholtz",
                                     "poisson_fused", "poisson_mode0"};
const char* dns_profile_kernel_name(int i) {
    return (i >= 0 && i < int(sizeof(kKernelNames) / sizeof(kKernelNames[0]))) ? kKernelNames[i] : nullptr;
}

int dns_profile_step(dns_ctx* c, int nsteps, double* out) {
    if (!c || !out || nsteps <= 0) return DNS_EINVAL;
    if (c->p.convection_scheme != 0 || c->p.pressure_solver != 0)
        return fail(c, DNS_EUNSUPPORTED, "dns_profile_step times the plain step's kernels only");
    const int NK = 7;
    int r = ensure_events(c, NK + 1);
    if (r) return r;
    std::vector<double> acc(NK, 0.0);
    for (int s = 0; s < nsteps; s++) {
        dns_begin_step(c);
        HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
        if ((r = nl_zderiv(c))) return r;
        HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
        if ((r = nl_c2r(c, false))) return r;
        HIPCHECK(c, hipEventRecord(c->ev[2], c->stream));
        if ((r = nl_r2c(c, false))) return r;
        HIPCHECK(c, hipEventRecord(c->ev[3], c->stream));
        // the fused step runs the flow controller inside the Helmholtz launch and mode (0,0)
        // inside the Poisson launch: their slots stay empty
        HIPCHECK(c, hipEventRecord(c->ev[4], c->stream));
        if ((r = viscous(c, true))) return r;
        HIPCHECK(c, hipEventRecord(c->ev[5], c->stream));
        if ((r = pressure(c, 1))) return r;
        if ((r = ctx_after_correction(c, c->stream))) return r;
        HIPCHECK(c, hipEventRecord(c->ev[6], c->stream));
        HIPCHECK(c, hipEventRecord(c->ev[7], c->stream));
        HIPCHECK(c, hipEventSynchronize(c->ev[7]));
        for (int i = 0; i < NK; i++) {
            float ms = 0.f;
            HIPCHECK(c, hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
            acc[i] += ms;
        }
    }
    for (int i = 0; i < NK; i++) out[i] = acc[i] / nsteps;
    return 0;
}

static int run_probe(double* out, hipError_t (*launch)(double*, hipStream_t)) {
    if (!out) return DNS_EINVAL;
    std::string why;
    if (!dns_check_gfx950(why)) { g_create_error = why; return DNS_EDEVICE; }
    double* d = nullptr;
    const size_t n = size_t(4096) * 256;
    if (hipMalloc(&d, n * sizeof(double)) != hipSuccess) return DNS_ENOMEM;
    hipError_t e = launch(d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out, d, n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? 0 : DNS_EDEVICE;
}

int dns_debug_copy(size_t n, int repeats, double* ms_out) {
    if (n < 2 || (n & 1) || repeats < 1) return DNS_EINVAL;
    std::string why;
    if (!dns_check_gfx950(why)) { g_create_error = why; return DNS_EDEVICE; }
    double *x = nullptr, *y = nullptr;
    if (hipMalloc(&x, n * sizeof(double)) != hipSuccess) return DNS_ENOMEM;
    if (hipMalloc(&y, n * sizeof(double)) != hipSuccess) { (void)hipFree(x); return DNS_ENOMEM; }
    hipError_t e = hipMemset(x, 0, n * sizeof(double));
    hipEvent_t a = nullptr, b = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, nullptr);
    for (int r = 0; r < repeats && e == hipSuccess; r++) e = launch_copy(x, y, n, nullptr);
    if (e == hipSuccess) e = hipEventRecord(b, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (ms_out) *ms_out = double(ms) / repeats;
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(x);
    (void)hipFree(y);
    return e == hipSuccess ? 0 : DNS_EDEVICE;
}

int dns_debug_mfma_probe(double* out) { return run_probe(out, launch_mfma_probe); }
int dns_debug_mfma_probe_f32(double* out) { return run_probe(out, launch_mfma_probe_f32); }

int dns_debug_initial_field(const dns_params* q, double* u, double* v, double* w) {
    if (!q || !u || !v || !w) return DNS_EINVAL;
    const std::string e = check_params(q);
    if (!e.empty()) { g_create_error = e; return DNS_EINVAL; }
    InitSpec s{q->nx, q->ny, q->nz, q->xlen, q->ylen, q->u_centerline, q->enable_perturbations,
               q->perturbation_kind, q->analytical_mode_x, q->analytical_mode_y, q->perturbation_amplitude,
               (uint64_t)q->seed, q->noise_kmax};
    std::vector<double> a, b, c;
    initial_field(s, a, b, c);
    std::memcpy(u, a.data(), a.size() * sizeof(double));
    std::memcpy(v, b.data(), b.size() * sizeof(double));
    std::memcpy(w, c.data(), c.size() * sizeof(double));
    return 0;
}

// dense (rows0 + rows1) x K: group-0 rows, then group-1 rows
static void unswizzle(const FoldedA& a, std::vector<double>& m) {
    const int rows = a.rows0 + a.rows1;
    m.assign(size_t(rows) * a.K, 0.0);
    auto get = [&](int cb, int r, int c) {
        const int s = c / 4, lane = (r % 16) + 16 * (c % 4);
        return a.data[(size_t(cb) * a.ksp + s) * 64 + lane];
    };
    for (int r = 0; r < a.rows0; r++)
        for (int c = 0; c < a.K; c++)
            m[size_t(r) * a.K + c] = (a.extra && r == a.rows0 - 1) ? a.data[size_t(2 * a.pb) * a.ksp * 64 + c]
                                                                   : get(r / 16, r, c);
    for (int r = 0; r < a.rows1; r++)
        for (int c = 0; c < a.K; c++) m[size_t(a.rows0 + r) * a.K + c] = get(a.pb + r / 16, r, c);
}

int dns_debug_wavenumbers(int nx, int ny, double xlen, double ylen, double* xw, double* yw) {
    if (nx < 2 || ny < 2 || !xw || !yw || !(xlen > 0.0) || !(ylen > 0.0)) return DNS_EINVAL;
    std::vector<double> x, y;
    wavenumbers(nx, ny, xlen, ylen, x, y);
    std::memcpy(xw, x.data(), x.size() * sizeof(double));
    std::memcpy(yw, y.data(), y.size() * sizeof(double));
    return 0;
}

int dns_debug_dealias_mask(int nx, int ny, int which, unsigned char* mask) {
    if (nx < 2 || ny < 2 || !mask || (which != 0 && which != 1)) return DNS_EINVAL;
    const int nxh = nx / 2 + 1;
    if (which == 0) {
        std::vector<unsigned char> m;
        dealias_mask(nx, ny, m);
        std::memcpy(mask, m.data(), m.size());
        return 0;
    }
    Dealias d{true, 0, 0, ny};
    dealias_bounds(nx, ny, d.kxm, d.kym);
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nxh; i++) mask[i + size_t(nxh) * j] = d.keep(i, j) ? 1 : 0;
    return 0;
}

int dns_debug_jacgs_stats(dns_ctx* c, unsigned long long* iterations, int* nonconverged) {
    if (!c || !iterations || !nonconverged) return DNS_EINVAL;
    if (c->p.pressure_solver != 1) return fail(c, DNS_EINVAL, "the context does not use the jacgs pressure solver");
    HIPCHECK(c, hipMemcpyAsync(iterations, c->jac_iters, sizeof(*iterations), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(nonconverged, c->jac_nonconv, sizeof(*nonconverged), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_debug_context_wavenumbers(dns_ctx* c, double* xw, double* yw) {
    if (!c || !xw || !yw) return DNS_EINVAL;
    HIPCHECK(c, hipMemcpyAsync(xw, c->xw, size_t(c->nxh) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(yw, c->yw, size_t(c->ny) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dns_debug_zmap(int nx, int ny, int* ncols, int* so, int* s2o, double* s2s) {
    if (nx < 2 || ny < 1 || nx % 2 || !ncols || !so || !s2o || !s2s) return DNS_EINVAL;
    const ZMap z = zmap_make(nx, ny, true);
    std::vector<int> t(size_t(8) * z.P), b(zmap_blocks(z));
    if (!zmap_build(z, t.data(), b.data())) return DNS_EINVAL;
    *ncols = 2 * z.P;
    for (int c = 0; c < 2 * z.P; c++) {
        so[c] = t[4 * c];
        s2o[c] = t[4 * c + 1];
        s2s[c] = s2o[c] < 0 ? 0.0 : s2o[c] == so[c] + 1 ? 0.0 : (so[c] & 1) ? -1.0 : 1.0;
    }
    return 0;
}

int dns_debug_operator(int nz, double dt, double re, int cn, int exact, int which, double* out, int* rows, int* cols) {
    Operators op;
    std::string err;
    if (!build_operators(nz, dt, re, cn ? 0.5 : 1.0, cn != 0, exact != 0, false, op, err)) {
        g_create_error = err;
        return DNS_ESOLVER;
    }
    std::vector<double> m;
    int r = 0, cc = 0;
    auto vec = [&](const std::vector<double>& v) { m = v; r = int(v.size()); cc = 1; };
    auto mat = [&](const FoldedA& a) { unswizzle(a, m); r = a.rows0 + a.rows1; cc = a.K; };
    switch (which) {
        case 0: {   // back to the logical column order [T | f T D2_I] from the kernel's panel order
            mat(op.hF);
            const std::vector<double> st = m;
            const int L1 = (cc - 1) / 2;
            for (int i = 0; i < r; i++)
                for (int p = 0; p < cc; p++) m[size_t(i) * cc + helm_panel_column(p, L1)] = st[size_t(i) * cc + p];
            break;
        }
        case 1: mat(op.hB); break;
        case 2: vec(op.hc1); break;
        case 3: vec(op.hc2); break;
        case 4: vec(op.hl); break;
        case 5: mat(op.pF); break;
        case 6: mat(op.pB); break;
        case 7: mat(op.pDB); break;
        case 8: vec(op.pl); break;
        case 9: m = op.m0; r = cc = nz - 1; break;
        case 10: mat(op.zD1); break;
        case 11: vec(op.wz); break;
        case 12:
            m.resize(size_t(nz) * nz);
            for (int i = 0; i < nz; i++)
                for (int j = 0; j < nz; j++) m[size_t(i) * nz + j] = op.d2[i + size_t(nz) * j];
            r = cc = nz;
            break;
        default: return DNS_EINVAL;
    }
    if (rows) *rows = r;
    if (cols) *cols = cc;
    if (out) std::memcpy(out, m.data(), m.size() * sizeof(double));
    return 0;
}

int dns_debug_operator_error(dns_ctx* c, double* out2) {
    if (!c || !out2) return DNS_EINVAL;
    out2[0] = c->op.helm_recon_err;
    out2[1] = c->op.pois_recon_err;
    return 0;
}

}  // extern "C"
