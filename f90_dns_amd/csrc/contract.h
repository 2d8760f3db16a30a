// contract.h -- block-level FP64 MFMA contraction of a 32-column tile of spectral columns.
//
// A block of 4 wavefronts owns 32 real columns (16 complex modes).  Output row blocks are split
// across the waves (wave w takes row blocks w, w+4, w+8, ...), so each row of the operator is
// read from HBM/L2 once per block; the B operand (the spectral data, prologue applied) is
// produced once per element by one thread and staged through double-buffered LDS chunks of 16
// rows, the next chunk prefetched into registers while the current one is consumed.
#pragma once
#include "device_common.h"

namespace dns {

constexpr int CT = 2;              // 16-column MFMA tiles per block
constexpr int BCOLS = 16 * CT;     // real columns per block
constexpr int BMODES = BCOLS / 2;  // complex modes per block
constexpr int KC = 16;             // B rows per LDS chunk (4 MFMA k-steps)
constexpr int LDB = BCOLS + 2;     // padded LDS row stride, doubles

// acc[i][ct] += A[rb = wave + 4i] x B, B produced by bprod(k, mode_local) -> double2 (Re, Im)
// of row k for the block's mode mode_local (zero beyond the operand's rows).  lds: 2*KC*LDB.
// Both operands are fetched one chunk ahead: B into registers then LDS, A (this wave's rows of
// the next four k-steps) into registers.
template <int RBW, class BP>
__device__ __forceinline__ void gemm_staged(d4 (&acc)[RBW][CT], const double* __restrict__ Asw, int RB, int KS,
                                            BP&& bprod, double* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kk = lane >> 4, cl = lane & 15;
    const int t = threadIdx.x;
    const int rt = t >> 4, mt = t & 15;            // staging: chunk row rt, mode mt
    const int nchunks = (KS + 3) / 4;
    double areg[RBW][4], anext[RBW][4];
    auto load_a = [&](double (&dst)[RBW][4], int ch) {
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            const int rb = wave + 4 * i;
#pragma unroll
            for (int sl = 0; sl < 4; sl++) {
                const int s = ch * 4 + sl;
                dst[i][sl] = (rb < RB && s < KS) ? Asw[(size_t(rb) * KS + s) * 64 + lane] : 0.0;
            }
        }
    };
    load_a(areg, 0);
    double2 v = bprod(rt, mt);
    lds[rt * LDB + 2 * mt] = v.x;
    lds[rt * LDB + 2 * mt + 1] = v.y;
    __syncthreads();
    for (int ch = 0; ch < nchunks; ch++) {
        const double* cur = lds + (ch & 1) * KC * LDB;
        double2 nxt = make_double2(0.0, 0.0);
        const bool more = ch + 1 < nchunks;
        if (more) {
            nxt = bprod((ch + 1) * KC + rt, mt);
            load_a(anext, ch + 1);
        }
#pragma unroll
        for (int sl = 0; sl < 4; sl++) {
            double b[CT];
#pragma unroll
            for (int ct = 0; ct < CT; ct++) b[ct] = cur[(sl * 4 + kk) * LDB + ct * 16 + cl];
#pragma unroll
            for (int i = 0; i < RBW; i++) {
                if (wave + 4 * i < RB) {
#pragma unroll
                    for (int ct = 0; ct < CT; ct++)
                        acc[i][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(areg[i][sl], b[ct], acc[i][ct], 0, 0, 0);
                }
            }
        }
        if (more) {
            double* nb = lds + ((ch + 1) & 1) * KC * LDB;
            nb[rt * LDB + 2 * mt] = nxt.x;
            nb[rt * LDB + 2 * mt + 1] = nxt.y;
#pragma unroll
            for (int i = 0; i < RBW; i++)
#pragma unroll
                for (int sl = 0; sl < 4; sl++) areg[i][sl] = anext[i][sl];
        }
        __syncthreads();
    }
}

// acc[i][ct] += A[rb = wave + 4i] x Z with Z resident in LDS as zs[row * LDB + col] (rows >= 4*KS
// never read).
template <int RBW>
__device__ __forceinline__ void gemm_lds(d4 (&acc)[RBW][CT], const double* __restrict__ Asw, int RB, int KS,
                                         const double* zs) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kk = lane >> 4, cl = lane & 15;
    double areg[RBW], anext[RBW];
#pragma unroll
    for (int i = 0; i < RBW; i++) areg[i] = (wave + 4 * i < RB) ? Asw[(size_t(wave + 4 * i) * KS) * 64 + lane] : 0.0;
    for (int s = 0; s < KS; s++) {
        const bool more = s + 1 < KS;
#pragma unroll
        for (int i = 0; i < RBW; i++)
            anext[i] = (more && wave + 4 * i < RB) ? Asw[(size_t(wave + 4 * i) * KS + s + 1) * 64 + lane] : 0.0;
        double b[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ct++) b[ct] = zs[(4 * s + kk) * LDB + ct * 16 + cl];
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            if (wave + 4 * i < RB) {
#pragma unroll
                for (int ct = 0; ct < CT; ct++)
                    acc[i][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(areg[i], b[ct], acc[i][ct], 0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < RBW; i++) areg[i] = anext[i];
    }
}

template <int RBW>
__device__ __forceinline__ void zero_acc2(d4 (&acc)[RBW][CT]) {
#pragma unroll
    for (int i = 0; i < RBW; i++)
#pragma unroll
        for (int ct = 0; ct < CT; ct++) acc[i][ct] = d4{0.0, 0.0, 0.0, 0.0};
}

// Output coordinates of acc[i][ct][r] in the block's column tile cb: row = (wave + 4i)*16 +
// lane/16 + 4r, column = cb*BCOLS + ct*16 + lane%16.
__device__ __forceinline__ int out_row(int i, int r) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    return (wave + 4 * i) * 16 + (lane >> 4) + 4 * r;
}
__device__ __forceinline__ int out_col(int ct, int cb) { return cb * BCOLS + ct * 16 + (threadIdx.x & 15); }

}  // namespace dns
