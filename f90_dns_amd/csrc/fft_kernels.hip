// fft_kernels.hip -- host side of the plane transforms (size classes, stage plans, launch
// dispatch to the instantiating units fft_c2r64/32.hip and fft_r2c64/32.hip), and the two-pass
// FP64 transforms for planes beyond the LDS.
#include <algorithm>
#include <vector>

#include "fft_plane.h"

namespace dns {

// ---- host: classes and plans ----------------------------------------------------------------
// plane (nx/2) x ny in whole swizzle blocks, twiddles ny + nx, column-0 buffer ny (complex);
// kx, ky (real)
static size_t lds_bytes(int nx, int ny, bool f32) {
    const size_t t = f32 ? sizeof(float) : sizeof(double);
    return 2 * t * (size_t(plane_region(nx / 2, ny)) + 2 * size_t(ny) + size_t(nx)) + t * (size_t(nx / 2 + 1) + ny);
}

// classes in order of capacity (and of threads per block)
static const int CLS_ORDER[FFT_CLASSES] = {0, 1, 2, 5, 6, 3, 4};

// the smallest class that holds the plane
static int size_class(int nx, int ny, bool f32) {
    const int H = (nx / 2) * ny;
    if (lds_bytes(nx, ny, f32) > 163840) return -1;
    for (int c : CLS_ORDER)
        if (H <= CLS_MAXE[c] && (c != 4 || f32)) return c;
    return -1;
}

// Stages for length n over nb transforms in class cls: at least two (the first reads the
// source, the last writes the result), fewest, then the smallest largest radix, every stage
// within the class's per-thread register budget (radix 2 and 3 always are, so every 2^a 3^b
// length n >= 4 has a plan).  Stages run in increasing radix: the first stage, which reads
// HBM, then has the most butterflies in flight.
static bool make_plan(int n, int nb, int cls, bool f32, Plan& p) {
    p.n = n; p.ns = 0;
    if (cls < 0 || n < 1) return false;
    auto fits = [&](int R) {
        const long long items = (long long)nb * (n / R);
        return (items + CLS_NT[cls] - 1) / CLS_NT[cls] <= stage_maxb(cls, R, f32);
    };
    std::vector<int> chosen;
    std::vector<std::pair<int, std::vector<int>>> stack;   // (remaining length, radices so far)
    stack.push_back({n, {}});
    while (!stack.empty()) {
        const int m = stack.back().first;
        const std::vector<int> seq = stack.back().second;
        stack.pop_back();
        if (m == 1) {
            if (seq.size() < 2) continue;
            if (chosen.empty() || seq.size() < chosen.size() ||
                (seq.size() == chosen.size() && seq.front() < chosen.front()))
                chosen = seq;
            continue;
        }
        const int maxr = seq.empty() ? 8 : seq.back();    // non-increasing: each multiset once
        for (int R : {8, 4, 3, 2}) {
            if (R > maxr || m % R || !fits(R)) continue;
            std::vector<int> s2 = seq;
            s2.push_back(R);
            stack.push_back({m / R, s2});
        }
    }
    if (chosen.empty() || chosen.size() > 8) return false;
    std::reverse(chosen.begin(), chosen.end());
    p.ns = int(chosen.size());
    for (int i = 0; i < p.ns; i++) p.r[i] = chosen[i];
    return true;
}

bool fft_geom(int nx, int ny, FFTGeom& g) {
    g.nx = nx; g.ny = ny; g.nxh = nx / 2 + 1; g.M = nx / 2;
    g.cls64 = size_class(nx, ny, false);
    g.cls32 = size_class(nx, ny, true);
    // plans stay empty (ns = 0) when nx/2 or ny is not 2^a 3^b; the plane launchers refuse them
    for (int f = 0; f < 2; f++) {
        const int cls = f ? g.cls32 : g.cls64;
        if (!(make_plan(ny, g.M, cls, f != 0, g.py[f]) && make_plan(g.M, ny, cls, f != 0, g.px[f])))
            g.py[f].ns = g.px[f].ns = 0;
    }
    return g.cls32 >= 0;
}

size_t plane_lds_bytes(const FFTGeom& g, bool f32) {
    return lds_bytes(g.nx, g.ny, f32);
}

// ---------------------------------------------------------------------------------------
// Two-pass FP64 plane transforms, for planes whose half-complex plane does not fit the LDS
// (nx/2)*ny > 8192 in FP64 (192 x 192).  Field input/output and diagnostics use them; the step
// itself runs such planes in FP32.  Pass "rows": LROWS x-lines per block (packed half-length
// complex FFT plus the real/half-complex split); pass "cols": LCOLS kx-columns per block
// (complex FFT along y), both staged in LDS, through a scratch plane in HBM.
// ---------------------------------------------------------------------------------------
constexpr int LROWS = 16, LCOLS = 16, TP_NT = 256, TP_MAXE = LROWS * 257;

// forward (sign -1): real rows -> half-complex rows;  inverse (+1): half-complex -> real, / scale
template <int SIGN>
__global__ __launch_bounds__(TP_NT) void k_tp_rows(TwoPassArgs a) {
    using V = double2;
    extern __shared__ char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    const V* twx = reinterpret_cast<const V*>(a.twx);
    const int p = blockIdx.y, y0 = blockIdx.x * LROWS;
    const int nxh = a.nxh, M = a.M, nx = a.nx;
    const int nr = min(LROWS, a.ny - y0);
    const int half = M / 2 + 1;
    if (SIGN < 0) {
        const double* src = a.src + size_t(p) * a.sstride + size_t(y0) * nx;
        for (int e = threadIdx.x; e < nr * M; e += TP_NT) {
            const int y = e / M, n = e - y * M;
            lds[y * nxh + n] = make_double2(src[size_t(y) * nx + 2 * n], src[size_t(y) * nx + 2 * n + 1]);
        }
        __syncthreads();
        fft_lds<TP_NT, TP_MAXE, false, V>(lds, a.fx, nr, 1, nxh, twx, nx, -1);
        for (int e = threadIdx.x; e < nr * half; e += TP_NT) {
            const int y = e / half, kk = e - y * half;
            V* row = lds + y * nxh;
            if (kk == 0) {
                const V Z0 = row[0];
                row[0] = make_double2(Z0.x + Z0.y, 0.0);
                row[M] = make_double2(Z0.x - Z0.y, 0.0);
            } else {
                const int k2 = M - kk;
                const V A = row[kk], B = row[k2];
                const V w = twx[kk];
                const V sm = make_double2(A.x + B.x, A.y - B.y);
                const V d = make_double2(A.x - B.x, A.y + B.y);
                const V wd = cmul(w, d);
                const V Xk = make_double2(0.5 * (sm.x + wd.y), 0.5 * (sm.y - wd.x));
                if (k2 != kk) {
                    const V w2 = make_double2(-w.x, w.y);
                    const V s2 = make_double2(B.x + A.x, B.y - A.y);
                    const V d2 = make_double2(B.x - A.x, B.y + A.y);
                    const V wd2 = cmul(w2, d2);
                    row[k2] = make_double2(0.5 * (s2.x + wd2.y), 0.5 * (s2.y - wd2.x));
                }
                row[kk] = Xk;
            }
        }
        __syncthreads();
        V* dst = reinterpret_cast<V*>(a.dst + size_t(p) * a.dstride) + size_t(y0) * nxh;
        for (int e = threadIdx.x; e < nr * nxh; e += TP_NT) dst[e] = lds[e];
    } else {
        const V* src = reinterpret_cast<const V*>(a.src + size_t(p) * a.sstride) + size_t(y0) * nxh;
        for (int e = threadIdx.x; e < nr * nxh; e += TP_NT) lds[e] = src[e];
        __syncthreads();
        for (int e = threadIdx.x; e < nr * half; e += TP_NT) {
            const int y = e / half, kk = e - y * half;
            V* row = lds + y * nxh;
            if (kk == 0) {
                const double x0 = row[0].x, xm = row[M].x;
                row[0] = make_double2(x0 + xm, x0 - xm);
            } else {
                const int k2 = M - kk;
                const V A = row[kk], B = row[k2];
                const V tw = twx[kk];
                const V w = make_double2(tw.x, -tw.y);
                const V sm = make_double2(A.x + B.x, A.y - B.y);
                const V d = make_double2(A.x - B.x, A.y + B.y);
                const V wd = cmul(w, d);
                const V Zk = make_double2(sm.x - wd.y, sm.y + wd.x);
                if (k2 != kk) {
                    const V w2 = make_double2(-w.x, w.y);
                    const V s2 = make_double2(B.x + A.x, B.y - A.y);
                    const V d2 = make_double2(B.x - A.x, B.y + A.y);
                    const V wd2 = cmul(w2, d2);
                    row[k2] = make_double2(s2.x - wd2.y, s2.y + wd2.x);
                }
                row[kk] = Zk;
            }
        }
        __syncthreads();
        fft_lds<TP_NT, TP_MAXE, false, V>(lds, a.fx, nr, 1, nxh, twx, nx, +1);
        double* dst = a.dst + size_t(p) * a.dstride + size_t(y0) * nx;
        for (int e = threadIdx.x; e < nr * M; e += TP_NT) {
            const int y = e / M, n = e - y * M;
            const V z = lds[y * nxh + n];
            dst[size_t(y) * nx + 2 * n] = z.x / a.scale;
            dst[size_t(y) * nx + 2 * n + 1] = z.y / a.scale;
        }
    }
}

// complex transforms along y of LCOLS kx-columns: element (y, c) at lds[y * LCOLS + c]
template <int SIGN>
__global__ __launch_bounds__(TP_NT) void k_tp_cols(TwoPassArgs a) {
    using V = double2;
    extern __shared__ char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    const V* twy = reinterpret_cast<const V*>(a.twy);
    const int p = blockIdx.y, c0 = blockIdx.x * LCOLS;
    const int nxh = a.nxh, ny = a.ny;
    const int nc = min(LCOLS, nxh - c0);
    const V* src = reinterpret_cast<const V*>(a.src + size_t(p) * a.sstride);
    for (int e = threadIdx.x; e < ny * LCOLS; e += TP_NT) {
        const int y = e / LCOLS, c = e - y * LCOLS;
        lds[e] = c < nc ? src[size_t(y) * nxh + c0 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<TP_NT, 257 * LCOLS, true, V>(lds, a.fy, LCOLS, LCOLS, 1, twy, ny, SIGN);
    V* dst = reinterpret_cast<V*>(a.dst + size_t(p) * a.dstride);
    const unsigned char* mask = a.mask;
    for (int e = threadIdx.x; e < ny * LCOLS; e += TP_NT) {
        const int y = e / LCOLS, c = e - y * LCOLS;
        if (c >= nc) continue;
        V x = lds[e];
        if (mask && !mask[size_t(y) * nxh + c0 + c]) x = make_double2(0.0, 0.0);
        dst[size_t(y) * nxh + c0 + c] = x;
    }
}

// spectral divergence i kx u + i ky v + dw of nplanes planes (the input of C2R_DIV)
__global__ __launch_bounds__(256) void k_div_spec(int nxh, int ny, long long total, const double* xw, const double* yw,
                                                  const double2* u, const double2* v, const double2* dw, double2* out) {
    const long long H = (long long)nxh * ny;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const int m = int(e % H), j = m / nxh, i = m - j * nxh;
        const double kx = xw[i], ky = yw[j];
        const double2 a = u[e], b = v[e], c = dw[e];
        out[e] = make_double2(-kx * a.y + -ky * b.y + c.x, kx * a.x + ky * b.x + c.y);
    }
}

hipError_t launch_div_spec(const FFTGeom& g, const double* xw, const double* yw, const double* u, const double* v,
                           const double* dw, double* out, int nplanes, hipStream_t s) {
    const long long total = (long long)g.nxh * g.ny * nplanes;
    const int blocks = int(std::min<long long>((total + 255) / 256, 4096));
    hipLaunchKernelGGL(k_div_spec, dim3(blocks), dim3(256), 0, s, g.nxh, g.ny, total, xw, yw,
                       reinterpret_cast<const double2*>(u), reinterpret_cast<const double2*>(v),
                       reinterpret_cast<const double2*>(dw), reinterpret_cast<double2*>(out));
    return hipGetLastError();
}

hipError_t launch_two_pass(const FFTGeom& g, bool forward, const double* src, double* dst, double* scratch,
                           int nplanes, long long sstride, long long dstride, hipStream_t s) {
    if (nplanes <= 0) return hipSuccess;
    if (g.M > 256 || g.ny > 256) return hipErrorInvalidValue;
    TwoPassArgs a{};
    a.nx = g.nx; a.ny = g.ny; a.nxh = g.nxh; a.M = g.M; a.fx = g.fx; a.fy = g.fy;
    a.twx = g.twx; a.twy = g.twy; a.scale = double(g.nx) * double(g.ny);
    const long long H2 = 2LL * g.nxh * g.ny;
    const dim3 grows((g.ny + LROWS - 1) / LROWS, nplanes), gcols((g.nxh + LCOLS - 1) / LCOLS, nplanes);
    const size_t lrows = size_t(LROWS) * g.nxh * sizeof(double2), lcols = size_t(LCOLS) * g.ny * sizeof(double2);
    if (forward) {
        a.src = src; a.dst = scratch; a.sstride = sstride; a.dstride = H2;
        hipLaunchKernelGGL(k_tp_rows<-1>, grows, dim3(TP_NT), lrows, s, a);
        a.src = scratch; a.dst = dst; a.sstride = H2; a.dstride = dstride;
        hipLaunchKernelGGL(k_tp_cols<-1>, gcols, dim3(TP_NT), lcols, s, a);
    } else {
        a.src = src; a.dst = scratch; a.sstride = sstride; a.dstride = H2;
        hipLaunchKernelGGL(k_tp_cols<+1>, gcols, dim3(TP_NT), lcols, s, a);
        a.src = scratch; a.dst = dst; a.sstride = H2; a.dstride = dstride;
        hipLaunchKernelGGL(k_tp_rows<+1>, grows, dim3(TP_NT), lrows, s, a);
    }
    return hipGetLastError();
}

static hipError_t prepare_two_pass() {
    const void* fns[] = {reinterpret_cast<const void*>(&k_tp_rows<-1>), reinterpret_cast<const void*>(&k_tp_rows<1>),
                         reinterpret_cast<const void*>(&k_tp_cols<-1>), reinterpret_cast<const void*>(&k_tp_cols<1>)};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t fft_prepare(const FFTGeom&) { return fft_prepare_all(); }

hipError_t fft_prepare_all() {
    static hipError_t done = [] {
        hipError_t e;
        if ((e = prepare_two_pass()) != hipSuccess || (e = c2r_prepare64()) != hipSuccess ||
            (e = c2r_prepare32()) != hipSuccess || (e = r2c_prepare64()) != hipSuccess ||
            (e = r2c_prepare32()) != hipSuccess)
            return e;
        return hipSuccess;
    }();
    return done;
}

int fft_class_capacity(int c, bool f32) {
    if (c < 0 || c >= FFT_CLASSES || (c == 4 && !f32)) return -1;
    return CLS_MAXE[c];
}

int fft_class_promote(int c, int levels, bool f32) {
    int i = 0;
    while (i < FFT_CLASSES && CLS_ORDER[i] != c) i++;
    if (i == FFT_CLASSES || fft_class_capacity(c, f32) < 0) return -1;
    i += levels;
    if (i >= FFT_CLASSES || fft_class_capacity(CLS_ORDER[i], f32) < 0) return -1;
    return CLS_ORDER[i];
}

void fft_fill(const FFTGeom& g, FFTArgs& a, bool f32) {
    a.nx = g.nx; a.ny = g.ny; a.nxh = g.nxh; a.M = g.M;
    // bounds of the 2/3 rule (the same as setup.cpp dealias_mask's table)
    dealias_bounds(g.nx, g.ny, a.kxm, a.kym);
    a.py = g.py[f32 ? 1 : 0]; a.px = g.px[f32 ? 1 : 0];
    if (f32) { a.twx = g.twx32; a.twy = g.twy32; }
    else { a.twx = g.twx; a.twy = g.twy; }
}

// Operand precisions the instances are built for: FP64 compute FP64 in and out; FP32 compute
// of the step NL (u, v, w FP64, omega_x, omega_y FP64 or FP32 per field), FP32 out; FP32
// CROSS FP32 in and out; FP32 PLAIN FP64 in and out; DIV FP64 only.
static bool operands_ok(bool c2r, int op, const FFTArgs& a, bool f32) {
    if (!f32) return a.src_f32 == 0 && a.dst_f32 == 0;
    if (c2r) {
        if (op == C2R_NL) return a.dst_f32 && (a.src_f32 & 7) == 0;
        if (op == C2R_PLAIN) return a.src_f32 == 0 && a.dst_f32 == 0;
        return false;
    }
    if (op == R2C_CROSS) return (a.src_f32 & 63) == 63 && a.dst_f32;
    return a.src_f32 == 0 && a.dst_f32 == 0;
}

static hipError_t dispatch(const FFTGeom& g, bool c2r, int op, FFTArgs a, int np, int nf, bool f32, hipStream_t s) {
    if (np <= 0 || nf <= 0) return hipSuccess;
    fft_fill(g, a, f32);
    const int cls = f32 ? g.cls32 : g.cls64;
    if (cls < 0 || a.py.ns < 2 || a.px.ns < 2 || !operands_ok(c2r, op, a, f32)) return hipErrorInvalidValue;
    // omega_z and the divergence read the wavenumber tables; every source and destination is read
    // or written
    if (c2r && op != C2R_PLAIN && (!a.xw || !a.yw)) return hipErrorInvalidValue;
    const int nsrc = c2r ? (op == C2R_NL ? 5 : op == C2R_DIV ? 3 : nf) : (op == R2C_CROSS ? 6 : nf);
    for (int q = 0; q < nsrc; q++)
        if (!a.src[q]) return hipErrorInvalidValue;
    for (int q = 0; q < nf; q++)
        if (!a.dst[q]) return hipErrorInvalidValue;
    const size_t lds = plane_lds_bytes(g, f32);
    if (c2r) return f32 ? c2r_launch32(cls, op, a, np, nf, lds, s) : c2r_launch64(cls, op, a, np, nf, lds, s);
    return f32 ? r2c_launch32(cls, op, a, np, nf, lds, s) : r2c_launch64(cls, op, a, np, nf, lds, s);
}

hipError_t launch_c2r(const FFTGeom& g, int op, FFTArgs a, int np, int nf, bool f32, hipStream_t s) {
    return dispatch(g, true, op, a, np, nf, f32, s);
}
hipError_t launch_r2c(const FFTGeom& g, int op, FFTArgs a, int np, int nf, bool f32, hipStream_t s) {
    return dispatch(g, false, op, a, np, nf, f32, s);
}

// The ragged launches are used by the fused step only (C2R_NL, R2C_CROSS); the descriptors
// are built by the caller with fft_fill and the step's operand precisions.
hipError_t launch_c2r_rag(int cls, int op, bool f32, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds,
                          hipStream_t s) {
    if (nt <= 0 || nf <= 0) return hipSuccess;
    if (op != C2R_NL) return hipErrorInvalidValue;
    return f32 ? c2r_launch_rag32(cls, d, map, nt, nf, lds, s) : c2r_launch_rag64(cls, d, map, nt, nf, lds, s);
}
hipError_t launch_r2c_rag(int cls, int op, bool f32, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds,
                          hipStream_t s) {
    if (nt <= 0 || nf <= 0) return hipSuccess;
    if (op != R2C_CROSS) return hipErrorInvalidValue;
    return f32 ? r2c_launch_rag32(cls, d, map, nt, nf, lds, s) : r2c_launch_rag64(cls, d, map, nt, nf, lds, s);
}

}  // namespace dns
