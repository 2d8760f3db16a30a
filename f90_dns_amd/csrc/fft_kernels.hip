// fft_kernels.hip -- 2D plane transforms real(nx,ny) <-> complex(nx/2+1,ny) with the whole
// half-complex plane resident in LDS (one workgroup per plane and field), in FP64 or FP32.
//
// c2r: complex inverse along y for every kx column, then the x direction as a half-length
// complex inverse of packed pairs z_n = x_{2n} + i x_{2n+1}; the packing uses only Re of the
// kx = 0 and kx = nx/2 entries, which is FFTW's c2r result for non-Hermitian input
// (fftw3_dns_module.f90:116-132).  Output divided by nx*ny.
// r2c: the mirror image, unnormalised (fftw3_dns_module.f90:95-108).
#include <algorithm>

#include "kernels.h"

namespace dns {

// plane-size classes: (threads, max complex elements of the half plane).  Class 0 is one
// wavefront per plane (its barriers are nearly free; small planes of ensembles keep many planes
// in flight per CU); class 4 only fits the LDS in FP32 (192 x 192).
template <int C> struct Cls;
template <> struct Cls<0> { static constexpr int NT = 64, MAXE = 1088; };
template <> struct Cls<1> { static constexpr int NT = 256, MAXE = 2176; };
template <> struct Cls<2> { static constexpr int NT = 512, MAXE = 4704; };
template <> struct Cls<3> { static constexpr int NT = 1024, MAXE = 8320; };
template <> struct Cls<4> { static constexpr int NT = 1024, MAXE = 18624; };

template <class T> struct Vec;
template <> struct Vec<double> { using type = double2; };
template <> struct Vec<float> { using type = float2; };

// load the complex (or real pair) at element index idx2 (in pairs) of a double or float array
template <class V>
__device__ __forceinline__ V ld2(const void* p, size_t idx2, bool f32) {
    if (f32) {
        const float2 x = reinterpret_cast<const float2*>(p)[idx2];
        return mk2(typename Scalar<V>::type(x.x), typename Scalar<V>::type(x.y));
    }
    const double2 x = reinterpret_cast<const double2*>(p)[idx2];
    return mk2(typename Scalar<V>::type(x.x), typename Scalar<V>::type(x.y));
}
template <class V>
__device__ __forceinline__ void st2(void* p, size_t idx2, V v, bool f32) {
    if (f32) reinterpret_cast<float2*>(p)[idx2] = make_float2(float(v.x), float(v.y));
    else reinterpret_cast<double2*>(p)[idx2] = make_double2(double(v.x), double(v.y));
}

// spectral input of c2r field f at plane k, mode (i, j), flat index e = i + nxh*j
template <int OP, class V>
__device__ __forceinline__ V c2r_input(const FFTArgs& a, int f, int k, int e, int i, int j) {
    using S = typename Scalar<V>::type;
    const size_t off2 = (size_t(k) * a.sstride) / 2 + size_t(e);
    if (OP == C2R_PLAIN) {
        return ld2<V>(a.src[f], off2, (a.src_f32 >> f) & 1);
    } else if (OP == C2R_NL) {
        // 0 u, 1 v, 2 w, 3 omega_x, 4 omega_y (precomputed), 5 omega_z = i kx v - i ky u
        if (f < 5) return ld2<V>(a.src[f], off2, (a.src_f32 >> f) & 1);
        const V u = ld2<V>(a.src[0], off2, a.src_f32 & 1);
        const V v = ld2<V>(a.src[1], off2, (a.src_f32 >> 1) & 1);
        const S kx = S(a.xw[i]), ky = S(a.yw[j]);
        // (i kx v) - (i ky u)
        return mk2(-kx * v.y - (-ky * u.y), kx * v.x - ky * u.x);
    } else {
        // divergence: i kx u + i ky v + (D1 w)
        const V u = ld2<V>(a.src[0], off2, a.src_f32 & 1);
        const V v = ld2<V>(a.src[1], off2, (a.src_f32 >> 1) & 1);
        const V dw = ld2<V>(a.src[2], off2, (a.src_f32 >> 2) & 1);
        const S kx = S(a.xw[i]), ky = S(a.yw[j]);
        return mk2(-kx * u.y + -ky * v.y + dw.x, kx * u.x + ky * v.x + dw.y);
    }
}

// plane k of field f
template <int C, int OP, class T>
__device__ __forceinline__ void c2r_plane(const FFTArgs& a, int k, int f) {
    using V = typename Vec<T>::type;
    constexpr int NT = Cls<C>::NT, MAXE = Cls<C>::MAXE;
    extern __shared__ char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    const V* twx = reinterpret_cast<const V*>(a.twx);
    const V* twy = reinterpret_cast<const V*>(a.twy);
    const int nxh = a.nxh, ny = a.ny, M = a.M, nx = a.nx;
    const int H = nxh * ny;
    for (int e = threadIdx.x; e < H; e += NT) {
        const int j = e / nxh, i = e - j * nxh;
        lds[e] = c2r_input<OP, V>(a, f, k, e, i, j);
    }
    __syncthreads();
    // inverse along y: nxh transforms of length ny, element stride nxh
    fft_lds<NT, MAXE, true, V>(lds, a.fy, nxh, nxh, 1, twy, ny, +1);
    // pack each row for the half-length inverse along x
    const int half = M / 2 + 1;
    for (int e = threadIdx.x; e < ny * half; e += NT) {
        const int y = e / half, kk = e - y * half;
        V* row = lds + y * nxh;
        if (kk == 0) {
            const T x0 = row[0].x, xm = row[M].x;
            row[0] = mk2(x0 + xm, x0 - xm);
        } else {
            const int k2 = M - kk;
            const V A = row[kk], B = row[k2];
            const V tw = twx[kk];
            const V w = mk2(tw.x, -tw.y);                    // exp(+2 pi i kk/nx)
            const V sm = mk2(A.x + B.x, A.y - B.y);          // A + conj B
            const V d = mk2(A.x - B.x, A.y + B.y);           // A - conj B
            const V wd = cmul(w, d);
            const V Zk = mk2(sm.x - wd.y, sm.y + wd.x);      // s + i w d
            if (k2 != kk) {
                const V w2 = mk2(-w.x, w.y);                 // exp(+2 pi i (M-kk)/nx)
                const V s2 = mk2(B.x + A.x, B.y - A.y);
                const V d2 = mk2(B.x - A.x, B.y + A.y);
                const V wd2 = cmul(w2, d2);
                row[k2] = mk2(s2.x - wd2.y, s2.y + wd2.x);
            }
            row[kk] = Zk;
        }
    }
    __syncthreads();
    // inverse along x: ny transforms of length M, contiguous, row stride nxh
    fft_lds<NT, MAXE, false, V>(lds, a.fx, ny, 1, nxh, twx, nx, +1);
    const T norm = T(nx) * T(ny);
    void* out = a.dst[f];
    const size_t ob2 = (size_t(k) * a.dstride) / 2;
    const bool of32 = a.dst_f32 != 0;
    for (int e = threadIdx.x; e < M * ny; e += NT) {
        const int y = e / M, n = e - y * M;
        const V z = lds[y * nxh + n];
        st2<V>(out, ob2 + size_t(y) * M + n, mk2(z.x / norm, z.y / norm), of32);
    }
}

template <int C, int OP, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_c2r(FFTArgs a) {
    c2r_plane<C, OP, T>(a, blockIdx.x, blockIdx.y);
}
template <int C, int OP, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_c2r_rag(const FFTArgs* __restrict__ d, const Tile* __restrict__ map) {
    const Tile t = map[blockIdx.x];
    const FFTArgs& a = d[t.g];
    c2r_plane<C, OP, T>(a, t.x, blockIdx.y);
}

template <int C, int OP, class T>
__device__ __forceinline__ void r2c_plane(const FFTArgs& a, int k, int f) {
    using V = typename Vec<T>::type;
    constexpr int NT = Cls<C>::NT, MAXE = Cls<C>::MAXE;
    extern __shared__ char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    const V* twx = reinterpret_cast<const V*>(a.twx);
    const V* twy = reinterpret_cast<const V*>(a.twy);
    const int nxh = a.nxh, ny = a.ny, M = a.M, nx = a.nx;
    const size_t po2 = (size_t(k) * a.sstride) / 2;
    for (int e = threadIdx.x; e < M * ny; e += NT) {
        const int y = e / M, n = e - y * M;
        const size_t q2 = po2 + size_t(y) * M + n;
        V z;
        if (OP == R2C_PLAIN) {
            z = ld2<V>(a.src[f], q2, (a.src_f32 >> f) & 1);
        } else {
            // rotational product S = u x omega (DNS_pressure_BC_3D.f90:1089-1091);
            // src = u, v, w, omega_x, omega_y, omega_z
            int ia, ib, ic, id;
            if (f == 0) { ia = 1; ib = 5; ic = 2; id = 4; }        // v*oz - w*oy
            else if (f == 1) { ia = 2; ib = 3; ic = 0; id = 5; }   // w*ox - u*oz
            else { ia = 0; ib = 4; ic = 1; id = 3; }               // u*oy - v*ox
            const V A = ld2<V>(a.src[ia], q2, (a.src_f32 >> ia) & 1);
            const V B = ld2<V>(a.src[ib], q2, (a.src_f32 >> ib) & 1);
            const V Cc = ld2<V>(a.src[ic], q2, (a.src_f32 >> ic) & 1);
            const V D = ld2<V>(a.src[id], q2, (a.src_f32 >> id) & 1);
            z = mk2(A.x * B.x - Cc.x * D.x, A.y * B.y - Cc.y * D.y);
        }
        lds[y * nxh + n] = z;
    }
    __syncthreads();
    // forward along x: half-length transforms of the packed rows
    fft_lds<NT, MAXE, false, V>(lds, a.fx, ny, 1, nxh, twx, nx, -1);
    const int half = M / 2 + 1;
    for (int e = threadIdx.x; e < ny * half; e += NT) {
        const int y = e / half, kk = e - y * half;
        V* row = lds + y * nxh;
        if (kk == 0) {
            const V Z0 = row[0];
            row[0] = mk2(Z0.x + Z0.y, T(0));
            row[M] = mk2(Z0.x - Z0.y, T(0));
        } else {
            const int k2 = M - kk;
            const V A = row[kk], B = row[k2];
            const V w = twx[kk];                                 // exp(-2 pi i kk/nx)
            const V sm = mk2(A.x + B.x, A.y - B.y);
            const V d = mk2(A.x - B.x, A.y + B.y);
            const V wd = cmul(w, d);
            const T h = T(0.5);
            const V Xk = mk2(h * (sm.x + wd.y), h * (sm.y - wd.x));   // (s - i w d)/2
            if (k2 != kk) {
                const V w2 = mk2(-w.x, w.y);                     // exp(-2 pi i (M-kk)/nx)
                const V s2 = mk2(B.x + A.x, B.y - A.y);
                const V d2 = mk2(B.x - A.x, B.y + A.y);
                const V wd2 = cmul(w2, d2);
                row[k2] = mk2(h * (s2.x + wd2.y), h * (s2.y - wd2.x));
            }
            row[kk] = Xk;
        }
    }
    __syncthreads();
    // forward along y
    fft_lds<NT, MAXE, true, V>(lds, a.fy, nxh, nxh, 1, twy, ny, -1);
    const int H = nxh * ny;
    void* out = a.dst[f];
    const size_t ob2 = (size_t(k) * a.dstride) / 2;
    const bool of32 = a.dst_f32 != 0;
    const unsigned char* mask = a.mask;
    for (int e = threadIdx.x; e < H; e += NT) {
        V x = lds[e];
        if (mask && !mask[e]) x = mk2(T(0), T(0));
        st2<V>(out, ob2 + e, x, of32);
    }
}

template <int C, int OP, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_r2c(FFTArgs a) {
    r2c_plane<C, OP, T>(a, blockIdx.x, blockIdx.y);
}
template <int C, int OP, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_r2c_rag(const FFTArgs* __restrict__ d, const Tile* __restrict__ map) {
    const Tile t = map[blockIdx.x];
    const FFTArgs& a = d[t.g];
    r2c_plane<C, OP, T>(a, t.x, blockIdx.y);
}

static int size_class(int H, bool f32) {
    if (H <= Cls<0>::MAXE) return 0;
    if (H <= Cls<1>::MAXE) return 1;
    if (H <= Cls<2>::MAXE) return 2;
    if (H <= Cls<3>::MAXE) return 3;
    if (f32 && H <= Cls<4>::MAXE) return 4;
    return -1;
}

bool fft_geom(int nx, int ny, FFTGeom& g) {
    g.nx = nx; g.ny = ny; g.nxh = nx / 2 + 1; g.M = nx / 2;
    const int H = g.nxh * ny;
    g.cls64 = size_class(H, false);
    g.cls32 = size_class(H, true);
    return g.cls32 >= 0;
}

// ---------------------------------------------------------------------------------------
// Two-pass FP64 plane transforms, for planes whose half-complex plane does not fit the LDS
// (nx/2+1)*ny > 8320 (192 x 192).  Field input/output and diagnostics use them; the step
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

template <int C, class T>
static hipError_t prepare_ct() {
    const void* fns[] = {reinterpret_cast<const void*>(&k_c2r<C, C2R_PLAIN, T>),
                         reinterpret_cast<const void*>(&k_c2r<C, C2R_NL, T>),
                         reinterpret_cast<const void*>(&k_c2r<C, C2R_DIV, T>),
                         reinterpret_cast<const void*>(&k_r2c<C, R2C_PLAIN, T>),
                         reinterpret_cast<const void*>(&k_r2c<C, R2C_CROSS, T>),
                         reinterpret_cast<const void*>(&k_c2r_rag<C, C2R_NL, T>),
                         reinterpret_cast<const void*>(&k_r2c_rag<C, R2C_CROSS, T>)};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
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

hipError_t fft_prepare(const FFTGeom& g) {
    hipError_t e = prepare_two_pass();
    if (e != hipSuccess) return e;
    switch (g.cls64) {
        case 0: e = prepare_ct<0, double>(); break;
        case 1: e = prepare_ct<1, double>(); break;
        case 2: e = prepare_ct<2, double>(); break;
        case 3: e = prepare_ct<3, double>(); break;
        default: break;
    }
    if (e != hipSuccess) return e;
    switch (g.cls32) {
        case 0: return prepare_ct<0, float>();
        case 1: return prepare_ct<1, float>();
        case 2: return prepare_ct<2, float>();
        case 3: return prepare_ct<3, float>();
        case 4: return prepare_ct<4, float>();
        default: return hipErrorInvalidValue;
    }
}

hipError_t fft_prepare_all() {
    hipError_t e;
    if ((e = prepare_two_pass()) != hipSuccess) return e;
    if ((e = prepare_ct<0, double>()) != hipSuccess || (e = prepare_ct<1, double>()) != hipSuccess ||
        (e = prepare_ct<2, double>()) != hipSuccess || (e = prepare_ct<3, double>()) != hipSuccess ||
        (e = prepare_ct<0, float>()) != hipSuccess || (e = prepare_ct<1, float>()) != hipSuccess ||
        (e = prepare_ct<2, float>()) != hipSuccess || (e = prepare_ct<3, float>()) != hipSuccess ||
        (e = prepare_ct<4, float>()) != hipSuccess)
        return e;
    return hipSuccess;
}

int fft_class_capacity(int c, bool f32) {
    switch (c) {
        case 0: return Cls<0>::MAXE;
        case 1: return Cls<1>::MAXE;
        case 2: return Cls<2>::MAXE;
        case 3: return Cls<3>::MAXE;
        case 4: return f32 ? Cls<4>::MAXE : -1;
        default: return -1;
    }
}

void fft_fill(const FFTGeom& g, FFTArgs& a, bool f32) {
    a.nx = g.nx; a.ny = g.ny; a.nxh = g.nxh; a.M = g.M;
    a.fx = g.fx; a.fy = g.fy;
    if (f32) { a.twx = g.twx32; a.twy = g.twy32; }
    else { a.twx = g.twx; a.twy = g.twy; }
}

template <int C, class T>
static hipError_t run(bool c2r, int op, const FFTArgs& a, int np, int nf, hipStream_t s) {
    const size_t lds = size_t(a.nxh) * a.ny * 2 * sizeof(T);
    const dim3 grid(np, nf), block(Cls<C>::NT);
    if (c2r) {
        if (op == C2R_PLAIN) hipLaunchKernelGGL((k_c2r<C, C2R_PLAIN, T>), grid, block, lds, s, a);
        else if (op == C2R_NL) hipLaunchKernelGGL((k_c2r<C, C2R_NL, T>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_c2r<C, C2R_DIV, T>), grid, block, lds, s, a);
    } else {
        if (op == R2C_PLAIN) hipLaunchKernelGGL((k_r2c<C, R2C_PLAIN, T>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_r2c<C, R2C_CROSS, T>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}

static hipError_t dispatch(const FFTGeom& g, bool c2r, int op, FFTArgs a, int np, int nf, bool f32, hipStream_t s) {
    if (np <= 0 || nf <= 0) return hipSuccess;
    fft_fill(g, a, f32);
    if (f32) {
        switch (g.cls32) {
            case 0: return run<0, float>(c2r, op, a, np, nf, s);
            case 1: return run<1, float>(c2r, op, a, np, nf, s);
            case 2: return run<2, float>(c2r, op, a, np, nf, s);
            case 3: return run<3, float>(c2r, op, a, np, nf, s);
            case 4: return run<4, float>(c2r, op, a, np, nf, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (g.cls64) {
        case 0: return run<0, double>(c2r, op, a, np, nf, s);
        case 1: return run<1, double>(c2r, op, a, np, nf, s);
        case 2: return run<2, double>(c2r, op, a, np, nf, s);
        case 3: return run<3, double>(c2r, op, a, np, nf, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_c2r(const FFTGeom& g, int op, FFTArgs a, int np, int nf, bool f32, hipStream_t s) {
    return dispatch(g, true, op, a, np, nf, f32, s);
}
hipError_t launch_r2c(const FFTGeom& g, int op, FFTArgs a, int np, int nf, bool f32, hipStream_t s) {
    return dispatch(g, false, op, a, np, nf, f32, s);
}

// The ragged launches are used by the fused step only (C2R_NL, R2C_CROSS).
template <int C, class T>
static hipError_t run_rag(bool c2r, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds, hipStream_t s) {
    const dim3 grid(nt, nf), block(Cls<C>::NT);
    if (c2r) hipLaunchKernelGGL((k_c2r_rag<C, C2R_NL, T>), grid, block, lds, s, d, map);
    else hipLaunchKernelGGL((k_r2c_rag<C, R2C_CROSS, T>), grid, block, lds, s, d, map);
    return hipGetLastError();
}

static hipError_t dispatch_rag(bool c2r, int cls, int op, bool f32, const FFTArgs* d, const Tile* map, int nt, int nf,
                               size_t lds, hipStream_t s) {
    if (nt <= 0 || nf <= 0) return hipSuccess;
    if (op != (c2r ? int(C2R_NL) : int(R2C_CROSS))) return hipErrorInvalidValue;
    if (f32) {
        switch (cls) {
            case 0: return run_rag<0, float>(c2r, d, map, nt, nf, lds, s);
            case 1: return run_rag<1, float>(c2r, d, map, nt, nf, lds, s);
            case 2: return run_rag<2, float>(c2r, d, map, nt, nf, lds, s);
            case 3: return run_rag<3, float>(c2r, d, map, nt, nf, lds, s);
            case 4: return run_rag<4, float>(c2r, d, map, nt, nf, lds, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (cls) {
        case 0: return run_rag<0, double>(c2r, d, map, nt, nf, lds, s);
        case 1: return run_rag<1, double>(c2r, d, map, nt, nf, lds, s);
        case 2: return run_rag<2, double>(c2r, d, map, nt, nf, lds, s);
        case 3: return run_rag<3, double>(c2r, d, map, nt, nf, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_c2r_rag(int cls, int op, bool f32, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds,
                          hipStream_t s) {
    return dispatch_rag(true, cls, op, f32, d, map, nt, nf, lds, s);
}
hipError_t launch_r2c_rag(int cls, int op, bool f32, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds,
                          hipStream_t s) {
    return dispatch_rag(false, cls, op, f32, d, map, nt, nf, lds, s);
}

}  // namespace dns
