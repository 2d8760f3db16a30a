// fft_plan.hip -- the FFT module of the C ABI (dns_fft_*): plane transforms on host or
// device arrays, the drop-in for fftw3_dns_module.f90:19-152.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/f90dns.h"
#include "kernels.h"
#include "setup.h"

using namespace dns;

// the context-creation error slot of dns_api.hip (dns_last_error(NULL))
void dns_set_create_error(const std::string& msg);
bool dns_check_gfx950(std::string& why);
bool dns_make_factors(int n, Factors& f);


// Scratch of dns_fft_execute_device_noalloc: the intermediate planes of the two-pass FP64
// transforms (planes that do not fit the LDS in FP64), in device memory that exists for the
// life of the process, one copy per device.  2 Mi doubles = 16 MiB: 15 planes of the largest
// supported size (512 x 256), 56 of 192 x 192; larger batches run in chunks of that many planes.
static constexpr size_t kScratchDoubles = size_t(2) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];
static constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices];   // device address of `scratch`, per device
static std::mutex g_scratch_mu;

// hipGetSymbolAddress of the scratch array on the current device, once per device (at plan
// creation); host code uses only these addresses, never the array's name
static int scratch_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        dns_set_create_error("device index beyond the scratch table");
        return DNS_EDEVICE;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[dev]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        dns_set_create_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return DNS_EDEVICE;
    }
    g_scratch_dev[dev] = static_cast<double*>(ptr);
    return 0;
}

struct dns_fft_plan {
    FFTGeom g;
    int dev = 0;                 // device the plan (its twiddles) lives on
    bool f32 = false;
    double *twx = nullptr, *twy = nullptr;
    float *twx32 = nullptr, *twy32 = nullptr;
    double *din = nullptr, *dout = nullptr;
    double* scr = nullptr;       // scratch of the two-pass FP64 transforms
    size_t cap_in = 0, cap_out = 0, cap_scr = 0;
    hipStream_t stream = nullptr;
};

static int plan_buf(double** p, size_t& cap, size_t n) {
    if (cap >= n) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr; cap = 0;
    if (hipMalloc(p, n * sizeof(double)) != hipSuccess) return DNS_ENOMEM;
    cap = n;
    return 0;
}

extern "C" {

int dns_fft_plan_create(int nx, int ny, int precision, dns_fft_plan** out) {
    if (!out) return DNS_EINVAL;
    *out = nullptr;
    if (precision != DNS_PREC_FP64 && precision != DNS_PREC_MIXED && precision != DNS_PREC_MIXED_F32D) {
        dns_set_create_error("bad precision");
        return DNS_EINVAL;
    }
    if (nx < 8 || ny < 4 || (nx & 1) || (ny & 1)) { dns_set_create_error("nx must be even and >= 8, ny even and >= 4"); return DNS_EINVAL; }
    std::string why;
    if (!dns_check_gfx950(why)) { dns_set_create_error(why); return DNS_EDEVICE; }
    dns_fft_plan* p = new dns_fft_plan();
    p->f32 = precision != DNS_PREC_FP64;
    if (hipGetDevice(&p->dev) != hipSuccess) { delete p; dns_set_create_error("hipGetDevice failed"); return DNS_EDEVICE; }
    if (int r = scratch_setup(p->dev)) { delete p; return r; }
    if (!fft_geom(nx, ny, p->g) || (!p->f32 && p->g.cls64 < 0 && (nx / 2 > 256 || ny > 256)) ||
        !dns_make_factors(ny, p->g.fy) ||
        !dns_make_factors(nx / 2, p->g.fx)) {
        delete p;
        dns_set_create_error("unsupported plane size");
        return DNS_EUNSUPPORTED;
    }
    std::vector<double> tw;
    twiddles(nx, tw);
    if (hipMalloc(&p->twx, tw.size() * sizeof(double)) != hipSuccess) { delete p; return DNS_ENOMEM; }
    (void)hipMemcpy(p->twx, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice);
    twiddles(ny, tw);
    if (hipMalloc(&p->twy, tw.size() * sizeof(double)) != hipSuccess) { (void)hipFree(p->twx); delete p; return DNS_ENOMEM; }
    (void)hipMemcpy(p->twy, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice);
    p->g.twx = reinterpret_cast<const double2*>(p->twx);
    p->g.twy = reinterpret_cast<const double2*>(p->twy);
    {
        std::vector<float> t32;
        twiddles(nx, tw); t32.assign(tw.begin(), tw.end());
        if (hipMalloc(&p->twx32, t32.size() * sizeof(float)) != hipSuccess) return DNS_ENOMEM;
        (void)hipMemcpy(p->twx32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice);
        twiddles(ny, tw); t32.assign(tw.begin(), tw.end());
        if (hipMalloc(&p->twy32, t32.size() * sizeof(float)) != hipSuccess) return DNS_ENOMEM;
        (void)hipMemcpy(p->twy32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice);
        p->g.twx32 = reinterpret_cast<const float2*>(p->twx32);
        p->g.twy32 = reinterpret_cast<const float2*>(p->twy32);
    }
    (void)fft_prepare(p->g);
    (void)hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    *out = p;
    return 0;
}

int dns_fft_plan_destroy(dns_fft_plan* p) {
    if (!p) return 0;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    (void)hipFree(p->twx); (void)hipFree(p->twy); (void)hipFree(p->twx32); (void)hipFree(p->twy32);
    if (p->din) (void)hipFree(p->din);
    if (p->dout) (void)hipFree(p->dout);
    if (p->scr) (void)hipFree(p->scr);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return 0;
}

int dns_fft_forward_device(dns_fft_plan* p, const double* in, double* out, int nplanes, void* stream) {
    if (!p || !in || !out || nplanes < 0) return DNS_EINVAL;
    FFTArgs a;
    std::memset(&a, 0, sizeof(a));
    a.src[0] = in; a.dst[0] = out;
    a.sstride = (long long)p->g.nx * p->g.ny; a.dstride = 2LL * p->g.nxh * p->g.ny;
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    if (!p->f32 && p->g.cls64 < 0) {
        if (plan_buf(&p->scr, p->cap_scr, size_t(a.dstride) * nplanes)) return DNS_ENOMEM;
        return launch_two_pass(p->g, true, in, out, p->scr, nplanes, a.sstride, a.dstride, s) == hipSuccess ? 0
                                                                                                       : DNS_EDEVICE;
    }
    return launch_r2c(p->g, R2C_PLAIN, a, nplanes, 1, p->f32, s) == hipSuccess ? 0 : DNS_EDEVICE;
}

int dns_fft_backward_device(dns_fft_plan* p, const double* in, double* out, int nplanes, void* stream) {
    if (!p || !in || !out || nplanes < 0) return DNS_EINVAL;
    FFTArgs a;
    std::memset(&a, 0, sizeof(a));
    a.src[0] = in; a.dst[0] = out;
    a.sstride = 2LL * p->g.nxh * p->g.ny; a.dstride = (long long)p->g.nx * p->g.ny;
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    if (!p->f32 && p->g.cls64 < 0) {
        if (plan_buf(&p->scr, p->cap_scr, size_t(a.sstride) * nplanes)) return DNS_ENOMEM;
        return launch_two_pass(p->g, false, in, out, p->scr, nplanes, a.sstride, a.dstride, s) == hipSuccess ? 0
                                                                                                        : DNS_EDEVICE;
    }
    return launch_c2r(p->g, C2R_PLAIN, a, nplanes, 1, p->f32, s) == hipSuccess ? 0 : DNS_EDEVICE;
}

int dns_fft_execute_device_noalloc(dns_fft_plan* p, int direction, const double* in, double* out, int nplanes,
                                   void* stream) {
    if (!p || !in || !out || nplanes < 0 || (direction != DNS_FFT_FORWARD && direction != DNS_FFT_BACKWARD)) {
        dns_set_create_error("dns_fft_execute_device_noalloc: bad argument");
        return DNS_EINVAL;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != p->dev) {
        dns_set_create_error("dns_fft_execute_device_noalloc: the plan's device is not the current device");
        return DNS_EINVAL;
    }
    const bool fwd = direction == DNS_FFT_FORWARD;
    const long long R = (long long)p->g.nx * p->g.ny, C = 2LL * p->g.nxh * p->g.ny;
    const long long sstride = fwd ? R : C, dstride = fwd ? C : R;
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    if (p->f32 || p->g.cls64 >= 0) {   // LDS-resident planes: no scratch
        FFTArgs a;
        std::memset(&a, 0, sizeof(a));
        a.src[0] = in; a.dst[0] = out;
        a.sstride = sstride; a.dstride = dstride;
        const hipError_t e = fwd ? launch_r2c(p->g, R2C_PLAIN, a, nplanes, 1, p->f32, s)
                                 : launch_c2r(p->g, C2R_PLAIN, a, nplanes, 1, p->f32, s);
        return e == hipSuccess ? 0 : DNS_EDEVICE;
    }
    // two passes through the scratch: one spectral plane (C doubles) per transformed plane
    if (size_t(C) > kScratchDoubles) {
        dns_set_create_error("dns_fft_execute_device_noalloc: one plane needs " + std::to_string(C) +
                             " doubles of scratch, more than the " + std::to_string(kScratchDoubles) + " it has");
        return DNS_EINVAL;
    }
    double* scr = g_scratch_dev[dev];
    if (!scr) {
        dns_set_create_error("dns_fft_execute_device_noalloc: scratch not set up on this device");
        return DNS_EDEVICE;
    }
    const int chunk = int(kScratchDoubles / size_t(C));
    for (int k0 = 0; k0 < nplanes; k0 += chunk) {
        const int n = std::min(chunk, nplanes - k0);
        if (launch_two_pass(p->g, fwd, in + sstride * k0, out + dstride * k0, scr, n, sstride, dstride, s) != hipSuccess)
            return DNS_EDEVICE;
    }
    return 0;
}

int dns_fft_forward(dns_fft_plan* p, const double* in, double* out, int nplanes) {
    if (!p || !in || !out || nplanes < 0) return DNS_EINVAL;
    size_t nr = size_t(p->g.nx) * p->g.ny * nplanes, nc = size_t(2) * p->g.nxh * p->g.ny * nplanes;
    if (plan_buf(&p->din, p->cap_in, nr) || plan_buf(&p->dout, p->cap_out, nc)) return DNS_ENOMEM;
    if (hipMemcpyAsync(p->din, in, nr * sizeof(double), hipMemcpyHostToDevice, p->stream) != hipSuccess) return DNS_EDEVICE;
    int r = dns_fft_forward_device(p, p->din, p->dout, nplanes, p->stream);
    if (r) return r;
    if (hipMemcpyAsync(out, p->dout, nc * sizeof(double), hipMemcpyDeviceToHost, p->stream) != hipSuccess) return DNS_EDEVICE;
    return hipStreamSynchronize(p->stream) == hipSuccess ? 0 : DNS_EDEVICE;
}

int dns_fft_backward(dns_fft_plan* p, const double* in, double* out, int nplanes) {
    if (!p || !in || !out || nplanes < 0) return DNS_EINVAL;
    size_t nr = size_t(p->g.nx) * p->g.ny * nplanes, nc = size_t(2) * p->g.nxh * p->g.ny * nplanes;
    if (plan_buf(&p->din, p->cap_in, nc) || plan_buf(&p->dout, p->cap_out, nr)) return DNS_ENOMEM;
    if (hipMemcpyAsync(p->din, in, nc * sizeof(double), hipMemcpyHostToDevice, p->stream) != hipSuccess) return DNS_EDEVICE;
    int r = dns_fft_backward_device(p, p->din, p->dout, nplanes, p->stream);
    if (r) return r;
    if (hipMemcpyAsync(out, p->dout, nr * sizeof(double), hipMemcpyDeviceToHost, p->stream) != hipSuccess) return DNS_EDEVICE;
    return hipStreamSynchronize(p->stream) == hipSuccess ? 0 : DNS_EDEVICE;
}

}  // extern "C"
