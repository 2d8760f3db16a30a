// fft_r2c64.hip -- FP64 instances of the r2c plane kernels: the step's R2C_CROSS in every
// FP64 size class (direct and ragged launches) and R2C_PLAIN in the largest class, which
// runs any plane of the precision (its plans fit every larger class).
#include "fft_plane.h"

namespace dns {

template <int C>
static hipError_t hot(const FFTArgs& a, int np, int nf, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_r2c<C, R2C_CROSS, double>), dim3(xcd_grid(np, nf)), dim3(Cls<C>::NT), lds, s, a, np, nf);
    return hipGetLastError();
}
template <int C>
static hipError_t hot_rag(const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_r2c_rag<C, R2C_CROSS, double>), dim3(xcd_grid(nt, nf)), dim3(Cls<C>::NT), lds, s, d, map, nt, nf);
    return hipGetLastError();
}

hipError_t r2c_launch64(int cls, int op, const FFTArgs& a, int np, int nf, size_t lds, hipStream_t s) {
    if (op == R2C_PLAIN) {
        hipLaunchKernelGGL((k_r2c<3, R2C_PLAIN, double>), dim3(xcd_grid(np, nf)), dim3(Cls<3>::NT), lds, s, a, np, nf);
        return hipGetLastError();
    }
    if (op != R2C_CROSS) return hipErrorInvalidValue;
    switch (cls) {
        case 0: return hot<0>(a, np, nf, lds, s);
        case 1: return hot<1>(a, np, nf, lds, s);
        case 2: return hot<2>(a, np, nf, lds, s);
        case 3: return hot<3>(a, np, nf, lds, s);
        case 5: return hot<5>(a, np, nf, lds, s);
        case 6: return hot<6>(a, np, nf, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t r2c_launch_rag64(int cls, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds, hipStream_t s) {
    switch (cls) {
        case 0: return hot_rag<0>(d, map, nt, nf, lds, s);
        case 1: return hot_rag<1>(d, map, nt, nf, lds, s);
        case 2: return hot_rag<2>(d, map, nt, nf, lds, s);
        case 3: return hot_rag<3>(d, map, nt, nf, lds, s);
        case 5: return hot_rag<5>(d, map, nt, nf, lds, s);
        case 6: return hot_rag<6>(d, map, nt, nf, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t r2c_prepare64() {
    const void* fns[] = {
        reinterpret_cast<const void*>(&k_r2c<0, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c_rag<0, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c<1, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c_rag<1, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c<2, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c_rag<2, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c<3, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c_rag<3, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c<5, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c_rag<5, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c<6, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c_rag<6, R2C_CROSS, double>),
        reinterpret_cast<const void*>(&k_r2c<3, R2C_PLAIN, double>),
    };
    for (const void* f : fns) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace dns
