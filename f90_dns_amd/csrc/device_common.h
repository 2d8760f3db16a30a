// device_common.h -- shared device code: complex helpers, the LDS-resident mixed-radix
// Stockham FFT used by the plane kernels, and the FP64 MFMA column-tile contraction.
#pragma once
#include <hip/hip_runtime.h>

namespace dns {

// Radix factors of one transform length (4s, then 2, then 3), passed by value.
struct Factors {
    int n, nf;
    int f[14];
};

// A device pointer as a global-memory pointer.  A pointer the compiler cannot trace to a kernel
// argument (one read from an argument block in memory, as the ragged launches do) is generic to
// it, and it accesses memory through it with flat instructions; those count against the LDS
// wait counter too, so every wait for LDS data would also wait for all global reads in flight.
// The integer round trip starts the pointer afresh in the global address space (the argument
// blocks hold device allocations only), and the accesses become global loads and stores.
template <class T> __device__ __forceinline__ T* gmem(T* p) {
    return (T*)(__attribute__((address_space(1))) T*)(__UINTPTR_TYPE__)p;
}

// complex arithmetic on double2 / float2
template <class V> struct Scalar;
template <> struct Scalar<double2> { using type = double; };
template <> struct Scalar<float2> { using type = float; };

__device__ __forceinline__ double2 mk2(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ float2 mk2(float a, float b) { return make_float2(a, b); }
template <class V> __device__ __forceinline__ V cadd(V a, V b) { return mk2(a.x + b.x, a.y + b.y); }
template <class V> __device__ __forceinline__ V csub(V a, V b) { return mk2(a.x - b.x, a.y - b.y); }
template <class V> __device__ __forceinline__ V cmul(V a, V b) {
    return mk2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// sign * i * z
template <class V> __device__ __forceinline__ V cmuli(V z, int sign) {
    return sign > 0 ? mk2(-z.y, z.x) : mk2(z.y, -z.x);
}

// In-place R-point DFT out_q = sum_r v_r exp(sign 2 pi i q r / R), unnormalised.
template <int R, class V> struct DFT;
template <class V> struct DFT<2, V> {
    static __device__ __forceinline__ void run(V (&v)[2], int) {
        V a = v[0], b = v[1];
        v[0] = cadd(a, b); v[1] = csub(a, b);
    }
};
template <class V> struct DFT<4, V> {
    static __device__ __forceinline__ void run(V (&v)[4], int sign) {
        V s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        V s13 = cadd(v[1], v[3]), d13 = cmuli(csub(v[1], v[3]), sign);
        v[0] = cadd(s02, s13); v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13); v[3] = csub(d02, d13);
    }
};
template <class V> struct DFT<3, V> {
    static __device__ __forceinline__ void run(V (&v)[3], int sign) {
        using S = typename Scalar<V>::type;
        const S c = S(-0.5), s = S(0.86602540378443864676372317075293618);   // cos, sin(2 pi/3)
        V a = v[0], sum = cadd(v[1], v[2]), dif = csub(v[1], v[2]);
        V t = mk2(a.x + c * sum.x, a.y + c * sum.y);
        V r = cmuli(mk2(s * dif.x, s * dif.y), sign);
        v[0] = cadd(a, sum);
        v[1] = cadd(t, r);
        v[2] = csub(t, r);
    }
};

// One Stockham stage over nb independent length-N transforms held in LDS.  Element e of
// transform b lives at lds[b*bs + e*es].  Every thread first reads all inputs of its
// butterflies into registers, then (after a barrier) writes the outputs, so the stage
// runs in place.  tw is exp(-2 pi i k / L) for k < L with L a multiple of N.
// BFAST: consecutive threads take consecutive transforms (b), else consecutive j.
template <int R, int NT, int MAXE, bool BFAST, class V>
__device__ __forceinline__ void fft_stage(V* lds, int N, int Ns, int nb, int es, int bs,
                                          const V* __restrict__ tw, int L, int sign) {
    constexpr int MAXB = (MAXE / R + NT - 1) / NT;
    const int m = N / R;
    const int total = nb * m;
    V v[MAXB][R];
    int bidx[MAXB], jidx[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int t = threadIdx.x + q * NT;
        if (t < total) {
            int b, j;
            if (BFAST) { j = t / nb; b = t - j * nb; }
            else { b = t / m; j = t - b * m; }
            bidx[q] = b; jidx[q] = j;
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = lds[b * bs + (j + r * m) * es];
        }
    }
    __syncthreads();
    const int twstep = L / (Ns * R);
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int t = threadIdx.x + q * NT;
        if (t < total) {
            const int b = bidx[q], j = jidx[q];
            const int kq = j / Ns;
            const int k = j - kq * Ns;
            if (k) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    V w = tw[k * r * twstep];
                    if (sign > 0) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            DFT<R, V>::run(v[q], sign);
            const int base = kq * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[b * bs + (base + r * Ns) * es] = v[q][r];
        }
    }
    __syncthreads();
}

template <int NT, int MAXE, bool BFAST, class V>
__device__ __forceinline__ void fft_lds(V* lds, const Factors& fac, int nb, int es, int bs,
                                        const V* __restrict__ tw, int L, int sign) {
    int Ns = 1;
    for (int s = 0; s < fac.nf; s++) {
        const int R = fac.f[s];
        if (R == 4) fft_stage<4, NT, MAXE, BFAST, V>(lds, fac.n, Ns, nb, es, bs, tw, L, sign);
        else if (R == 2) fft_stage<2, NT, MAXE, BFAST, V>(lds, fac.n, Ns, nb, es, bs, tw, L, sign);
        else fft_stage<3, NT, MAXE, BFAST, V>(lds, fac.n, Ns, nb, es, bs, tw, L, sign);
        Ns *= R;
    }
}

// ---------------------------------------------------------------------------------------
// FP64 MFMA contraction of one wavefront's 16-column tile:
//   acc[rb] += A[rb*16 .. rb*16+15][0 .. 4*KS) * B[0 .. 4*KS)[c0 .. c0+15]
// A is pre-swizzled on the host to the A-operand lane order of v_mfma_f64_16x16x4f64
// (entry [(rb*KS + s)*64 + lane] = A(rb*16 + lane%16, 4s + lane/16)); bload(k) returns
// B(k, c0 + lane%16).  Result: acc[rb][r] = C(rb*16 + lane/16 + 4*r, c0 + lane%16) (the result
// layout of the instruction, checked by the lane probe in the GPU tests).
// ---------------------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));

template <int RB, class BL>
__device__ __forceinline__ void mfma_tile(d4 (&acc)[RB], const double* __restrict__ Asw, int KS, BL&& bload) {
    const int lane = threadIdx.x & 63;
    const int kk = lane >> 4;
    const double* ap = Asw + lane;
    for (int s = 0; s < KS; s++) {
        const double b = bload(4 * s + kk);
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const double a = ap[(size_t(rb) * KS + s) * 64];
            acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[rb], 0, 0, 0);
        }
    }
}

template <int RB>
__device__ __forceinline__ void zero_acc(d4 (&acc)[RB]) {
#pragma unroll
    for (int rb = 0; rb < RB; rb++) acc[rb] = d4{0.0, 0.0, 0.0, 0.0};
}

}  // namespace dns
