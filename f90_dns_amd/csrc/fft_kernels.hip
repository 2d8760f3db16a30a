// fft_kernels.hip -- 2D plane transforms real(nx,ny) <-> complex(nx/2+1,ny) with the whole
// half-complex plane resident in LDS (one workgroup per plane and field).
//
// c2r: complex inverse along y for every kx column, then the x direction as a half-length
// complex inverse of packed pairs z_n = x_{2n} + i x_{2n+1}; the packing uses only Re of the
// kx = 0 and kx = nx/2 entries, which is FFTW's c2r result for non-Hermitian input
// (fftw3_dns_module.f90:116-132).  Output divided by nx*ny.
// r2c: the mirror image, unnormalised (fftw3_dns_module.f90:95-108).
#include "kernels.h"

namespace dns {

// plane-size classes: (threads, max complex elements of the half plane)
template <int C> struct Cls;
template <> struct Cls<0> { static constexpr int NT = 256, MAXE = 2176; };
template <> struct Cls<1> { static constexpr int NT = 512, MAXE = 4704; };
template <> struct Cls<2> { static constexpr int NT = 1024, MAXE = 8320; };

// spectral input of c2r field f at plane k, mode (i, j), flat index e = i + nxh*j
template <int OP>
__device__ __forceinline__ double2 c2r_input(const FFTArgs& a, int f, int k, int e, int i, int j) {
    const size_t off = size_t(k) * a.sstride + 2 * size_t(e);
    if (OP == C2R_PLAIN) {
        return *reinterpret_cast<const double2*>(a.src[f] + off);
    } else if (OP == C2R_NL) {
        // 0 u, 1 v, 2 w, 3 omega_x, 4 omega_y (precomputed), 5 omega_z = i kx v - i ky u
        if (f < 5) return *reinterpret_cast<const double2*>(a.src[f] + off);
        const double2 u = *reinterpret_cast<const double2*>(a.src[0] + off);
        const double2 v = *reinterpret_cast<const double2*>(a.src[1] + off);
        const double kx = a.xw[i], ky = a.yw[j];
        // (i kx v) - (i ky u)
        return make_double2(-kx * v.y - (-ky * u.y), kx * v.x - ky * u.x);
    } else {
        // divergence: i kx u + i ky v + (D1 w)
        const double2 u = *reinterpret_cast<const double2*>(a.src[0] + off);
        const double2 v = *reinterpret_cast<const double2*>(a.src[1] + off);
        const double2 dw = *reinterpret_cast<const double2*>(a.src[2] + off);
        const double kx = a.xw[i], ky = a.yw[j];
        return make_double2(-kx * u.y + -ky * v.y + dw.x, kx * u.x + ky * v.x + dw.y);
    }
}

template <int C, int OP>
__global__ __launch_bounds__(Cls<C>::NT) void k_c2r(FFTArgs a) {
    constexpr int NT = Cls<C>::NT, MAXE = Cls<C>::MAXE;
    extern __shared__ double2 lds[];
    const int k = blockIdx.x, f = blockIdx.y;
    const int nxh = a.nxh, ny = a.ny, M = a.M, nx = a.nx;
    const int H = nxh * ny;
    for (int e = threadIdx.x; e < H; e += NT) {
        const int j = e / nxh, i = e - j * nxh;
        lds[e] = c2r_input<OP>(a, f, k, e, i, j);
    }
    __syncthreads();
    // inverse along y: nxh transforms of length ny, element stride nxh
    fft_lds<NT, MAXE, true>(lds, a.fy, nxh, nxh, 1, a.twy, ny, +1);
    // pack each row for the half-length inverse along x
    const int half = M / 2 + 1;
    for (int e = threadIdx.x; e < ny * half; e += NT) {
        const int y = e / half, kk = e - y * half;
        double2* row = lds + y * nxh;
        if (kk == 0) {
            const double x0 = row[0].x, xm = row[M].x;
            row[0] = make_double2(x0 + xm, x0 - xm);
        } else {
            const int k2 = M - kk;
            const double2 A = row[kk], B = row[k2];
            const double2 tw = a.twx[kk];
            const double2 w = make_double2(tw.x, -tw.y);                 // exp(+2 pi i kk/nx)
            const double2 s = make_double2(A.x + B.x, A.y - B.y);        // A + conj B
            const double2 d = make_double2(A.x - B.x, A.y + B.y);        // A - conj B
            const double2 wd = cmul(w, d);
            const double2 Zk = make_double2(s.x - wd.y, s.y + wd.x);     // s + i w d
            if (k2 != kk) {
                const double2 w2 = make_double2(-w.x, w.y);              // exp(+2 pi i (M-kk)/nx)
                const double2 s2 = make_double2(B.x + A.x, B.y - A.y);
                const double2 d2 = make_double2(B.x - A.x, B.y + A.y);
                const double2 wd2 = cmul(w2, d2);
                row[k2] = make_double2(s2.x - wd2.y, s2.y + wd2.x);
            }
            row[kk] = Zk;
        }
    }
    __syncthreads();
    // inverse along x: ny transforms of length M, contiguous, row stride nxh
    fft_lds<NT, MAXE, false>(lds, a.fx, ny, 1, nxh, a.twx, nx, +1);
    const double norm = double(nx) * double(ny);
    double* out = a.dst[f] + size_t(k) * a.dstride;
    for (int e = threadIdx.x; e < M * ny; e += NT) {
        const int y = e / M, n = e - y * M;
        const double2 z = lds[y * nxh + n];
        reinterpret_cast<double2*>(out + size_t(y) * nx)[n] = make_double2(z.x / norm, z.y / norm);
    }
}

template <int C, int OP>
__global__ __launch_bounds__(Cls<C>::NT) void k_r2c(FFTArgs a) {
    constexpr int NT = Cls<C>::NT, MAXE = Cls<C>::MAXE;
    extern __shared__ double2 lds[];
    const int k = blockIdx.x, f = blockIdx.y;
    const int nxh = a.nxh, ny = a.ny, M = a.M, nx = a.nx;
    const size_t po = size_t(k) * a.sstride;
    for (int e = threadIdx.x; e < M * ny; e += NT) {
        const int y = e / M, n = e - y * M;
        const size_t q = po + size_t(y) * nx + 2 * size_t(n);
        double2 z;
        if (OP == R2C_PLAIN) {
            z = *reinterpret_cast<const double2*>(a.src[f] + q);
        } else {
            // rotational product S = u x omega (DNS_pressure_BC_3D.f90:1089-1091);
            // src = u, v, w, omega_x, omega_y, omega_z
            int ia, ib, ic, id;
            if (f == 0) { ia = 1; ib = 5; ic = 2; id = 4; }        // v*oz - w*oy
            else if (f == 1) { ia = 2; ib = 3; ic = 0; id = 5; }   // w*ox - u*oz
            else { ia = 0; ib = 4; ic = 1; id = 3; }               // u*oy - v*ox
            const double2 A = *reinterpret_cast<const double2*>(a.src[ia] + q);
            const double2 B = *reinterpret_cast<const double2*>(a.src[ib] + q);
            const double2 Cc = *reinterpret_cast<const double2*>(a.src[ic] + q);
            const double2 D = *reinterpret_cast<const double2*>(a.src[id] + q);
            z = make_double2(A.x * B.x - Cc.x * D.x, A.y * B.y - Cc.y * D.y);
        }
        lds[y * nxh + n] = z;
    }
    __syncthreads();
    // forward along x: half-length transforms of the packed rows
    fft_lds<NT, MAXE, false>(lds, a.fx, ny, 1, nxh, a.twx, nx, -1);
    const int half = M / 2 + 1;
    for (int e = threadIdx.x; e < ny * half; e += NT) {
        const int y = e / half, kk = e - y * half;
        double2* row = lds + y * nxh;
        if (kk == 0) {
            const double2 Z0 = row[0];
            row[0] = make_double2(Z0.x + Z0.y, 0.0);
            row[M] = make_double2(Z0.x - Z0.y, 0.0);
        } else {
            const int k2 = M - kk;
            const double2 A = row[kk], B = row[k2];
            const double2 w = a.twx[kk];                                 // exp(-2 pi i kk/nx)
            const double2 s = make_double2(A.x + B.x, A.y - B.y);
            const double2 d = make_double2(A.x - B.x, A.y + B.y);
            const double2 wd = cmul(w, d);
            const double2 Xk = make_double2(0.5 * (s.x + wd.y), 0.5 * (s.y - wd.x));   // (s - i w d)/2
            if (k2 != kk) {
                const double2 w2 = make_double2(-w.x, w.y);              // exp(-2 pi i (M-kk)/nx)
                const double2 s2 = make_double2(B.x + A.x, B.y - A.y);
                const double2 d2 = make_double2(B.x - A.x, B.y + A.y);
                const double2 wd2 = cmul(w2, d2);
                row[k2] = make_double2(0.5 * (s2.x + wd2.y), 0.5 * (s2.y - wd2.x));
            }
            row[kk] = Xk;
        }
    }
    __syncthreads();
    // forward along y
    fft_lds<NT, MAXE, true>(lds, a.fy, nxh, nxh, 1, a.twy, ny, -1);
    const int H = nxh * ny;
    double* out = a.dst[f] + size_t(k) * a.dstride;
    const unsigned char* mask = a.mask;
    for (int e = threadIdx.x; e < H; e += NT) {
        double2 x = lds[e];
        if (mask && !mask[e]) x = make_double2(0.0, 0.0);
        reinterpret_cast<double2*>(out)[e] = x;
    }
}

bool fft_geom(int nx, int ny, FFTGeom& g) {
    g.nx = nx; g.ny = ny; g.nxh = nx / 2 + 1; g.M = nx / 2;
    const int H = g.nxh * ny;
    g.lds_bytes = size_t(H) * sizeof(double2);
    if (H <= Cls<0>::MAXE) g.cls = 0;
    else if (H <= Cls<1>::MAXE) g.cls = 1;
    else if (H <= Cls<2>::MAXE) g.cls = 2;
    else g.cls = -1;
    return g.cls >= 0;
}

template <int C, int OP>
static hipError_t run_c2r(const FFTGeom& g, const FFTArgs& a, int np, int nf, hipStream_t s) {
    hipLaunchKernelGGL((k_c2r<C, OP>), dim3(np, nf), dim3(Cls<C>::NT), g.lds_bytes, s, a);
    return hipGetLastError();
}
template <int C, int OP>
static hipError_t run_r2c(const FFTGeom& g, const FFTArgs& a, int np, int nf, hipStream_t s) {
    hipLaunchKernelGGL((k_r2c<C, OP>), dim3(np, nf), dim3(Cls<C>::NT), g.lds_bytes, s, a);
    return hipGetLastError();
}

template <int C>
static hipError_t c2r_cls(const FFTGeom& g, int op, const FFTArgs& a, int np, int nf, hipStream_t s) {
    switch (op) {
        case C2R_PLAIN: return run_c2r<C, C2R_PLAIN>(g, a, np, nf, s);
        case C2R_NL: return run_c2r<C, C2R_NL>(g, a, np, nf, s);
        default: return run_c2r<C, C2R_DIV>(g, a, np, nf, s);
    }
}
template <int C>
static hipError_t r2c_cls(const FFTGeom& g, int op, const FFTArgs& a, int np, int nf, hipStream_t s) {
    if (op == R2C_PLAIN) return run_r2c<C, R2C_PLAIN>(g, a, np, nf, s);
    return run_r2c<C, R2C_CROSS>(g, a, np, nf, s);
}

template <int C>
static hipError_t prepare_cls() {
    const void* fns[] = {reinterpret_cast<const void*>(&k_c2r<C, C2R_PLAIN>), reinterpret_cast<const void*>(&k_c2r<C, C2R_NL>),
                         reinterpret_cast<const void*>(&k_c2r<C, C2R_DIV>), reinterpret_cast<const void*>(&k_r2c<C, R2C_PLAIN>),
                         reinterpret_cast<const void*>(&k_r2c<C, R2C_CROSS>)};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t fft_prepare(const FFTGeom& g) {
    switch (g.cls) {
        case 0: return prepare_cls<0>();
        case 1: return prepare_cls<1>();
        case 2: return prepare_cls<2>();
        default: return hipErrorInvalidValue;
    }
}

static void fill_geom(const FFTGeom& g, FFTArgs& a) {
    a.nx = g.nx; a.ny = g.ny; a.nxh = g.nxh; a.M = g.M;
    a.fx = g.fx; a.fy = g.fy; a.twx = g.twx; a.twy = g.twy;
}

hipError_t launch_c2r(const FFTGeom& g, int op, FFTArgs a, int np, int nf, hipStream_t s) {
    if (np <= 0 || nf <= 0) return hipSuccess;
    fill_geom(g, a);
    switch (g.cls) {
        case 0: return c2r_cls<0>(g, op, a, np, nf, s);
        case 1: return c2r_cls<1>(g, op, a, np, nf, s);
        case 2: return c2r_cls<2>(g, op, a, np, nf, s);
        default: return hipErrorInvalidValue;
    }
}
hipError_t launch_r2c(const FFTGeom& g, int op, FFTArgs a, int np, int nf, hipStream_t s) {
    if (np <= 0 || nf <= 0) return hipSuccess;
    fill_geom(g, a);
    switch (g.cls) {
        case 0: return r2c_cls<0>(g, op, a, np, nf, s);
        case 1: return r2c_cls<1>(g, op, a, np, nf, s);
        case 2: return r2c_cls<2>(g, op, a, np, nf, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace dns
