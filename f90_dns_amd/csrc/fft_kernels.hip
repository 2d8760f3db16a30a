// fft_kernels.hip -- 2D plane transforms real(nx,ny) <-> complex(nx/2+1,ny) with the whole
// half-complex plane resident in LDS (one workgroup per plane and field), in FP64 or FP32.
//
// c2r: complex inverse along y for every kx column, then the x direction as a half-length
// complex inverse of packed pairs z_n = x_{2n} + i x_{2n+1}; the packing uses only Re of the
// kx = 0 and kx = nx/2 entries, which is FFTW's c2r result for non-Hermitian input
// (fftw3_dns_module.f90:116-132).  Output divided by nx*ny.
// r2c: the mirror image, unnormalised (fftw3_dns_module.f90:95-108).
#include "kernels.h"

namespace dns {

// plane-size classes: (threads, max complex elements of the half plane); class 3 only fits
// the LDS in FP32 (192 x 192)
template <int C> struct Cls;
template <> struct Cls<0> { static constexpr int NT = 256, MAXE = 2176; };
template <> struct Cls<1> { static constexpr int NT = 512, MAXE = 4704; };
template <> struct Cls<2> { static constexpr int NT = 1024, MAXE = 8320; };
template <> struct Cls<3> { static constexpr int NT = 1024, MAXE = 18624; };

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

template <int C, int OP, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_c2r(FFTArgs a) {
    using V = typename Vec<T>::type;
    constexpr int NT = Cls<C>::NT, MAXE = Cls<C>::MAXE;
    extern __shared__ char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    const V* twx = reinterpret_cast<const V*>(a.twx);
    const V* twy = reinterpret_cast<const V*>(a.twy);
    const int k = blockIdx.x, f = blockIdx.y;
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
__global__ __launch_bounds__(Cls<C>::NT) void k_r2c(FFTArgs a) {
    using V = typename Vec<T>::type;
    constexpr int NT = Cls<C>::NT, MAXE = Cls<C>::MAXE;
    extern __shared__ char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    const V* twx = reinterpret_cast<const V*>(a.twx);
    const V* twy = reinterpret_cast<const V*>(a.twy);
    const int k = blockIdx.x, f = blockIdx.y;
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

static int size_class(int H, bool f32) {
    if (H <= Cls<0>::MAXE) return 0;
    if (H <= Cls<1>::MAXE) return 1;
    if (H <= Cls<2>::MAXE) return 2;
    if (f32 && H <= Cls<3>::MAXE) return 3;
    return -1;
}

bool fft_geom(int nx, int ny, FFTGeom& g) {
    g.nx = nx; g.ny = ny; g.nxh = nx / 2 + 1; g.M = nx / 2;
    const int H = g.nxh * ny;
    g.cls64 = size_class(H, false);
    g.cls32 = size_class(H, true);
    return g.cls32 >= 0;
}

template <int C, class T>
static hipError_t prepare_ct() {
    const void* fns[] = {reinterpret_cast<const void*>(&k_c2r<C, C2R_PLAIN, T>),
                         reinterpret_cast<const void*>(&k_c2r<C, C2R_NL, T>),
                         reinterpret_cast<const void*>(&k_c2r<C, C2R_DIV, T>),
                         reinterpret_cast<const void*>(&k_r2c<C, R2C_PLAIN, T>),
                         reinterpret_cast<const void*>(&k_r2c<C, R2C_CROSS, T>)};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t fft_prepare(const FFTGeom& g) {
    hipError_t e = hipSuccess;
    switch (g.cls64) {
        case 0: e = prepare_ct<0, double>(); break;
        case 1: e = prepare_ct<1, double>(); break;
        case 2: e = prepare_ct<2, double>(); break;
        default: break;
    }
    if (e != hipSuccess) return e;
    switch (g.cls32) {
        case 0: return prepare_ct<0, float>();
        case 1: return prepare_ct<1, float>();
        case 2: return prepare_ct<2, float>();
        case 3: return prepare_ct<3, float>();
        default: return hipErrorInvalidValue;
    }
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
    a.nx = g.nx; a.ny = g.ny; a.nxh = g.nxh; a.M = g.M;
    a.fx = g.fx; a.fy = g.fy;
    if (f32) {
        a.twx = g.twx32; a.twy = g.twy32;
        switch (g.cls32) {
            case 0: return run<0, float>(c2r, op, a, np, nf, s);
            case 1: return run<1, float>(c2r, op, a, np, nf, s);
            case 2: return run<2, float>(c2r, op, a, np, nf, s);
            case 3: return run<3, float>(c2r, op, a, np, nf, s);
            default: return hipErrorInvalidValue;
        }
    }
    a.twx = g.twx; a.twy = g.twy;
    switch (g.cls64) {
        case 0: return run<0, double>(c2r, op, a, np, nf, s);
        case 1: return run<1, double>(c2r, op, a, np, nf, s);
        case 2: return run<2, double>(c2r, op, a, np, nf, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_c2r(const FFTGeom& g, int op, FFTArgs a, int np, int nf, bool f32, hipStream_t s) {
    return dispatch(g, true, op, a, np, nf, f32, s);
}
hipError_t launch_r2c(const FFTGeom& g, int op, FFTArgs a, int np, int nf, bool f32, hipStream_t s) {
    return dispatch(g, false, op, a, np, nf, f32, s);
}

}  // namespace dns
