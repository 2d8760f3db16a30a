// fft_c2r32.hip -- FP32 instances of the c2r plane kernels: the step's C2R_NL in every
// FP32 size class (direct and ragged launches, persistent in class 3) and C2R_PLAIN in the
// largest class, which runs any plane of the precision (its plans fit every larger class).
#include <algorithm>

#include "fft_plane.h"

namespace dns {

template <int C>
static hipError_t hot(const FFTArgs& a, int np, int nf, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_c2r<C, C2R_NL, float>), dim3(xcd_grid(np, nf)), dim3(Cls<C>::NT), lds, s, a, np, nf);
    return hipGetLastError();
}
// The step's c2r of class 3 runs persistent, one block per CU (k_c2r_nl_pers), when its first y
// stage has a radix the kernel is built for; cus is set by c2r_prepare32.
#ifndef C2R_PERSIST
#define C2R_PERSIST 1
#endif
static int g_cus = 0;
template <int R0>
static hipError_t pers(const FFTArgs& a, int np, size_t lds, hipStream_t s) {
    const int nv = xcd_grid(np, 6);
    const int g = std::min(nv, (g_cus / XCDS) * XCDS);
    hipLaunchKernelGGL((k_c2r_nl_pers<3, R0, float>), dim3(g), dim3(Cls<3>::NT), lds, s, a, np, nv);
    return hipGetLastError();
}
static hipError_t hot3(const FFTArgs& a, int np, int nf, size_t lds, hipStream_t s) {
    if (C2R_PERSIST && nf == 6 && g_cus >= XCDS) {
        switch (a.py.r[0]) {
            case 8: return pers<8>(a, np, lds, s);
            case 4: return pers<4>(a, np, lds, s);
            case 3: return pers<3>(a, np, lds, s);
            case 2: return pers<2>(a, np, lds, s);
            default: break;
        }
    }
    return hot<3>(a, np, nf, lds, s);
}
template <int C>
static hipError_t hot_rag(const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_c2r_rag<C, C2R_NL, float>), dim3(xcd_grid(nt, nf)), dim3(Cls<C>::NT), lds, s, d, map, nt, nf);
    return hipGetLastError();
}

hipError_t c2r_launch32(int cls, int op, const FFTArgs& a, int np, int nf, size_t lds, hipStream_t s) {
    if (op == C2R_PLAIN) {
        hipLaunchKernelGGL((k_c2r<4, C2R_PLAIN, float>), dim3(xcd_grid(np, nf)), dim3(Cls<4>::NT), lds, s, a, np, nf);
        return hipGetLastError();
    }
    if (op != C2R_NL) return hipErrorInvalidValue;
    switch (cls) {
        case 0: return hot<0>(a, np, nf, lds, s);
        case 1: return hot<1>(a, np, nf, lds, s);
        case 2: return hot<2>(a, np, nf, lds, s);
        case 3: return hot3(a, np, nf, lds, s);
        case 4: return hot<4>(a, np, nf, lds, s);
        case 5: return hot<5>(a, np, nf, lds, s);
        case 6: return hot<6>(a, np, nf, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t c2r_launch_rag32(int cls, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds, hipStream_t s) {
    switch (cls) {
        case 0: return hot_rag<0>(d, map, nt, nf, lds, s);
        case 1: return hot_rag<1>(d, map, nt, nf, lds, s);
        case 2: return hot_rag<2>(d, map, nt, nf, lds, s);
        case 3: return hot_rag<3>(d, map, nt, nf, lds, s);
        case 4: return hot_rag<4>(d, map, nt, nf, lds, s);
        case 5: return hot_rag<5>(d, map, nt, nf, lds, s);
        case 6: return hot_rag<6>(d, map, nt, nf, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t c2r_prepare32() {
    const void* fns[] = {
        reinterpret_cast<const void*>(&k_c2r<0, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r_rag<0, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r<1, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r_rag<1, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r<2, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r_rag<2, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r<3, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r_rag<3, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r<4, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r_rag<4, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r<5, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r_rag<5, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r<6, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r_rag<6, C2R_NL, float>),
        reinterpret_cast<const void*>(&k_c2r<4, C2R_PLAIN, float>),
        reinterpret_cast<const void*>(&k_c2r_nl_pers<3, 8, float>),
        reinterpret_cast<const void*>(&k_c2r_nl_pers<3, 4, float>),
        reinterpret_cast<const void*>(&k_c2r_nl_pers<3, 3, float>),
        reinterpret_cast<const void*>(&k_c2r_nl_pers<3, 2, float>),
    };
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    for (const void* f : fns) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace dns
