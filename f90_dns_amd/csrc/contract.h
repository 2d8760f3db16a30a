// contract.h -- block-level MFMA contractions (FP64; FP32 for the vorticity's D1 in the
// FP32-contraction precision) of the folded wall-normal operators over a 32-column tile of
// spectral columns.
//
// A block of NW wavefronts (4; at small nz one per row-block pair) owns 32 real columns (16
// complex modes).  A folded operator has two
// row groups (setup.h FoldedA): group 0 acts on panel 0, group 1 on panel 1, both K rows.  Wave
// w takes the row-block pairs b = w, w + NW, ...: group-0 block b and group-1 block b, so the two
// results of one output row pair (g0_i, g1_i) sit in the same lane and item and unfold in
// registers.  The panels (the spectral data, folded and prologue applied) are produced once per
// element by one thread and staged through double-buffered LDS chunks of rows; the next
// chunk and the wave's A operands for it are fetched while the current one is consumed.
#pragma once
#include "device_common.h"
#include "kernels.h"

namespace dns {

constexpr int CT = 2;              // 16-column MFMA tiles per block
constexpr int BCOLS = 16 * CT;     // real columns per block
constexpr int BMODES = BCOLS / 2;  // complex modes per block
constexpr int KC = 16;             // panel rows per chunk of an LDS-resident contraction (4 k-steps)
constexpr int LDB = BCOLS + 2;     // padded LDS row stride, doubles
// k-steps per staged chunk of fold_gemm in a block of NW waves (with 8, the prefetch registers
// cost occupancy and the kernels measured slower; blocks of one or two waves take shorter chunks
// so that each thread stages one row of a chunk).  Each must divide 8: ksp is a multiple of 8,
// so the A operands of the chunk after the last one stay inside the block.
#ifndef FOLD_KS_NW1
#define FOLD_KS_NW1 1
#endif
#ifndef FOLD_KS_NW2
#define FOLD_KS_NW2 2
#endif
#ifndef FOLD_KS_NW3
#define FOLD_KS_NW3 2
#endif
__host__ __device__ constexpr int fold_chunk_ks(int nw) {
    return nw == 1 ? FOLD_KS_NW1 : nw == 2 ? FOLD_KS_NW2 : nw == 3 ? FOLD_KS_NW3 : 4;
}
static_assert(8 % fold_chunk_ks(1) == 0 && 8 % fold_chunk_ks(2) == 0 && 8 % fold_chunk_ks(3) == 0,
              "chunk k-steps must divide 8");
// chunks by which the panel reads of fold_gemm run ahead of the MFMAs (four waves;
// FOLD_PREFETCH_SMALL for fewer)
#ifndef FOLD_PREFETCH
#define FOLD_PREFETCH 1
#endif
#ifndef FOLD_PREFETCH_SMALL
#define FOLD_PREFETCH_SMALL 1
#endif
constexpr int STAGE_LDS = 4 * 4 * 4 * LDB;   // doubles of the staging buffers (2 buffers x 2 panels x 16 rows)
// the staging buffers of fold_gemm in a block of nw waves: 2 buffers x 2 panels x 4 fold_chunk_ks(nw)
// rows (blocks of fewer waves stage shorter chunks, so they need less LDS and more of them fit a CU)
__host__ __device__ constexpr int stage_lds(int nw) { return 4 * 4 * fold_chunk_ks(nw) * LDB; }
static_assert(stage_lds(4) == STAGE_LDS, "four-wave staging");
static_assert(stage_lds(1) >= 512 && stage_lds(2) >= 512 && stage_lds(3) >= 512, "reduce_extra's 16 x 16 partials");
// waves per block of a contraction kernel variant (z_variant): 1, 2, 3, then 4
__host__ __device__ constexpr int z_variant_nw(int var) { return var < 3 ? var + 1 : 4; }

// the element at byte offset boff of a uniform base: with the base in scalar registers and a
// 32-bit offset the load takes the scalar-base form (no per-lane 64-bit address arithmetic)
template <class T>
__device__ __forceinline__ T ld_at(const T* base, unsigned boff) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}

// Element type of a contraction.  FP64: v_mfma_f64_16x16x4f64; FP32: v_mfma_f32_16x16x4f32
// with FP32 accumulation.  Both take A(i, k) at lane i + 16 k and B(k, j) at lane j + 16 k (the
// stored operator layout); their results differ in layout (lane probes in the GPU tests): item
// r of lane l holds row l/16 + 4 r (FP64) or row 4 (l/16) + r (FP32) of the 16-row block, at
// column l % 16.
typedef float f4 __attribute__((ext_vector_type(4)));
template <class T> struct Mfma;
template <> struct Mfma<double> {
    using acc = d4;
    static __device__ __forceinline__ acc op(double a, double b, const acc& c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
    static __device__ __forceinline__ const double* A(const FoldOp& o) { return gmem(o.A); }
};
template <> struct Mfma<float> {
    using acc = f4;
    static __device__ __forceinline__ acc op(float a, float b, const acc& c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static __device__ __forceinline__ const float* A(const FoldOp& o) { return gmem(o.Af); }
};
template <class T> using acc_t = typename Mfma<T>::acc;

struct FoldPair {
    double2 p0, p1;   // (Re, Im) of the panel-0 and panel-1 entries of one row and mode
};

// A staging thread's partial sums of the dense extra row times panel 0 (fold_gemm), per mode.
// Blocks of one, two or four waves stage RS = 4 NW rows per pass, a divisor of 16: the partials
// are kept per row class kf mod 16 (16 / RS slots per thread, each summed in row order) and
// reduce_extra adds the 16 classes in order, so the extra row's value does not depend on the
// waves per block (a column block may run at any of those widths with the same results).  Three
// waves keep one partial per thread (their rows of a class fall to different threads).  Slot j
// holds class rt + RS ((base + j) mod N); fold_gemm rotates the slots between chunks so that
// each staged row adds into a slot known at compile time.
template <int NW> struct XAcc {
    static constexpr int RS = 4 * NW;
    static constexpr bool BYCLASS = 16 % RS == 0;
    static constexpr int N = BYCLASS ? 16 / RS : 1;
    double2 s[N];
    int base;
    __device__ __forceinline__ XAcc() : base(0) {
#pragma unroll
        for (int j = 0; j < N; j++) s[j] = make_double2(0.0, 0.0);
    }
};

// acc0[i][ct] += A0[block b] x P0, acc1[i][ct] += A1[block b] x P1 for b = wave + NW i (a
// block of NW waves).
// Panel row kf for the staging thread's mode (thread t of the block's NW waves stages rows
// t/16 + 4 NW r of a chunk for mode t%16) is finish(kf, fetch(kf)), zero beyond K: fetch issues the memory reads (branch
// free), finish does the arithmetic, so the next chunk's reads are in flight during the
// current chunk's MFMA chain.  lds: stage_lds(NW) doubles.
// With op.extra, xacc accumulates the staging thread's share of the dense extra row times
// panel 0 (XAcc; reduce_extra sums the shares of a column).
// SKIP: waves without row blocks skip the MFMA chain (operators with fewer row-block pairs
// than waves, the small-nz variants); otherwise every wave has blocks and the loop body stays
// one straight-line block.
// T: the element type of the panels, the A operands and the MFMAs (FP32 stages the folded
// values rounded once to FP32 and reads the operator's FP32 copy op.Af; the extra row stays FP64).
template <int RBW, bool SKIP, int NW, class T = double, class FETCH, class FINISH>
__device__ __forceinline__ void fold_gemm(acc_t<T> (&acc0)[RBW][CT], acc_t<T> (&acc1)[RBW][CT], const FoldOp& op,
                                          FETCH&& fetch, FINISH&& finish, T* lds, XAcc<NW>& xacc) {
    using Raw = decltype(fetch(0));
    constexpr int KS = fold_chunk_ks(NW);    // k-steps per chunk
    constexpr int KR = 4 * KS;               // panel rows per chunk
    constexpr int RS = 4 * NW;               // rows one pass of the block's threads stages
    constexpr int RPT = (KR + RS - 1) / RS;  // rows staged per thread (the last pass may be partial)
    const double* xrow = gmem(fold_xrow(op));
    const int xlast = 4 * op.ksp - 1;        // the extra row's stored entries (zero beyond K)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int kk = lane >> 4, cl = lane & 15;
    const int t = threadIdx.x, rt = t >> 4, mt = t & 15;
    const int nch = (op.ks + KS - 1) / KS;
    const bool active = wave < (op.rb0 > op.rb1 ? op.rb0 : op.rb1);
    // the wave's A blocks (uniform bases; element (k-step s, lane) at element 64 s + lane)
    const T* a0p[RBW];
    const T* a1p[RBW];
#pragma unroll
    for (int i = 0; i < RBW; i++) {
        const int b = wave + NW * i;
        a0p[i] = Mfma<T>::A(op) + (size_t(b) * op.ksp) * 64;
        a1p[i] = Mfma<T>::A(op) + (size_t(op.pb + b) * op.ksp) * 64;
    }
    auto aoff = [&](int s) { return unsigned((s * 64 + lane) * sizeof(T)); };
    // A operands of the current chunk; each is refilled with the next chunk's k-step as soon as
    // its MFMAs are issued (a rolling one-chunk prefetch in KS registers per group)
    T ar0[RBW][KS], ar1[RBW][KS];
#pragma unroll
    for (int i = 0; i < RBW; i++)
#pragma unroll
        for (int sl = 0; sl < KS; sl++) {
            ar0[i][sl] = ld_at(a0p[i], aoff(sl));
            ar1[i][sl] = ld_at(a1p[i], aoff(sl));
        }
    // the extra-row coefficient of a staged row (clamped: rows past K have zero panel entries)
    auto xcoef = [&](int kf) { return op.extra ? xrow[min(kf, xlast)] : 0.0; };
    // the extra-row partials by row class: row rt + RS rr of a chunk adds into slot rr mod N;
    // between chunks the slots rotate by the chunk's rows in units of RS
    using XA = XAcc<NW>;
    static_assert(!XA::BYCLASS || KR % RS == 0, "chunks of whole staging passes");
    constexpr int XSH = XA::BYCLASS ? (KR / RS) % XA::N : 0;
    auto xrotate = [&]() {
        if constexpr (XSH != 0) {
            double2 t[XA::N];
#pragma unroll
            for (int j = 0; j < XA::N; j++) t[j] = xacc.s[(j + XSH) % XA::N];
#pragma unroll
            for (int j = 0; j < XA::N; j++) xacc.s[j] = t[j];
            xacc.base = (xacc.base + XSH) % XA::N;
        }
    };
    auto stage = [&](T* buf, const FoldPair& v, int row, double c, double2& xs) {
        if (RS * RPT > KR && row >= KR) return;   // a partial last pass (wave-uniform)
        xs.x += c * v.p0.x;
        xs.y += c * v.p0.y;
        buf[row * LDB + 2 * mt] = T(v.p0.x);
        buf[row * LDB + 2 * mt + 1] = T(v.p0.y);
        buf[KR * LDB + row * LDB + 2 * mt] = T(v.p1.x);
        buf[KR * LDB + row * LDB + 2 * mt + 1] = T(v.p1.y);
    };
    // Panel reads run PD chunks ahead of the MFMAs and the arithmetic of the next chunk's fold
    // runs beside the current chunk's MFMA chain: the loop body is one straight-line block (a
    // chunk past the end folds rows past K, which are zero, into the idle buffer; every wave
    // issues the chain, absent row blocks having zero A operands).  q[d] holds the reads of
    // chunk ch + 1 + d.
    constexpr int PD = NW < 4 ? FOLD_PREFETCH_SMALL : FOLD_PREFETCH;
    Raw q[PD][RPT];
    double qc[PD][RPT];
    {
        Raw r0[RPT];
#pragma unroll
        for (int rr = 0; rr < RPT; rr++) {
            r0[rr] = fetch(rt + RS * rr);
#pragma unroll
            for (int d = 0; d < PD; d++) {
                q[d][rr] = fetch((1 + d) * KR + rt + RS * rr);
                qc[d][rr] = xcoef((1 + d) * KR + rt + RS * rr);
            }
        }
#pragma unroll
        for (int rr = 0; rr < RPT; rr++)
            stage(lds, finish(rt + RS * rr, r0[rr]), rt + RS * rr, xcoef(rt + RS * rr), xacc.s[rr % XA::N]);
    }
    __syncthreads();
    for (int ch = 0; ch < nch; ch++) {
        const T* cur = lds + (ch & 1) * 2 * KR * LDB;
        T* nb = lds + ((ch + 1) & 1) * 2 * KR * LDB;
        Raw r2[RPT];
        double c2[RPT];
        FoldPair f1[RPT];
#pragma unroll
        for (int rr = 0; rr < RPT; rr++) {
            const int kf2 = (ch + 1 + PD) * KR + rt + RS * rr;
            r2[rr] = fetch(kf2);
            c2[rr] = xcoef(kf2);
            f1[rr] = finish((ch + 1) * KR + rt + RS * rr, q[0][rr]);
        }
        const int nxa = min(ch + 1, nch - 1) * KS;   // the next chunk's A operands (the last chunk reloads its own)
        if (!SKIP || active)
#pragma unroll
        for (int sl = 0; sl < KS; sl++) {
            T b0[CT], b1[CT];
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                b0[ct] = cur[(sl * 4 + kk) * LDB + ct * 16 + cl];
                b1[ct] = cur[KR * LDB + (sl * 4 + kk) * LDB + ct * 16 + cl];
            }
#pragma unroll
            for (int i = 0; i < RBW; i++) {
#pragma unroll
                for (int ct = 0; ct < CT; ct++) {
                    acc0[i][ct] = Mfma<T>::op(ar0[i][sl], b0[ct], acc0[i][ct]);
                    acc1[i][ct] = Mfma<T>::op(ar1[i][sl], b1[ct], acc1[i][ct]);
                }
                ar0[i][sl] = ld_at(a0p[i], aoff(nxa + sl));
                ar1[i][sl] = ld_at(a1p[i], aoff(nxa + sl));
            }
        }
        xrotate();
#pragma unroll
        for (int rr = 0; rr < RPT; rr++) {
            stage(nb, f1[rr], rt + RS * rr, qc[0][rr], xacc.s[rr % XA::N]);
#pragma unroll
            for (int d = 0; d + 1 < PD; d++) {
                q[d][rr] = q[d + 1][rr];
                qc[d][rr] = qc[d + 1][rr];
            }
            q[PD - 1][rr] = r2[rr];
            qc[PD - 1][rr] = c2[rr];
        }
        __syncthreads();
    }
}

// The same contraction with both panels resident in LDS (z0, z1: >= 4*ks rows, row stride LDB);
// the A operands are fetched one chunk of 4 k-steps ahead (ksp is a multiple of 8 >= ks + 1,
// so the chunk after the last one stays inside the block).  The last chunk stops at k-step ks
// (the k-steps past it would add exact zeros: zero A columns on zero panel rows).
template <int RBW, int NW>
__device__ __forceinline__ void fold_gemm_lds(d4 (&acc0)[RBW][CT], d4 (&acc1)[RBW][CT], const FoldOp& op,
                                              const double* z0, const double* z1) {
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int kk = lane >> 4, cl = lane & 15;
    const bool active = wave < (op.rb0 > op.rb1 ? op.rb0 : op.rb1);
    if (!active) return;
    const int nch = (op.ks + 3) / 4;
    const double* a0p[RBW];
    const double* a1p[RBW];
    auto aoff = [&](int s) { return unsigned((s * 64 + lane) * 8); };
    double ar0[RBW][4], ar1[RBW][4], an0[RBW][4], an1[RBW][4];
#pragma unroll
    for (int i = 0; i < RBW; i++) {
        const int b = wave + NW * i;
        a0p[i] = gmem(op.A) + (size_t(b) * op.ksp) * 64;
        a1p[i] = gmem(op.A) + (size_t(op.pb + b) * op.ksp) * 64;
#pragma unroll
        for (int sl = 0; sl < 4; sl++) {
            ar0[i][sl] = ld_at(a0p[i], aoff(sl));
            ar1[i][sl] = ld_at(a1p[i], aoff(sl));
        }
    }
    for (int ch = 0; ch < nch; ch++) {
        if (ch + 1 < nch) {
#pragma unroll
            for (int i = 0; i < RBW; i++)
#pragma unroll
                for (int sl = 0; sl < 4; sl++) {
                    an0[i][sl] = ld_at(a0p[i], aoff((ch + 1) * 4 + sl));
                    an1[i][sl] = ld_at(a1p[i], aoff((ch + 1) * 4 + sl));
                }
        }
#pragma unroll
        for (int sl = 0; sl < 4; sl++) {
            const int s = ch * 4 + sl;
            if (s >= op.ks) break;    // wave-uniform: the partial last chunk
            double b0[CT], b1[CT];
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                b0[ct] = z0[(4 * s + kk) * LDB + ct * 16 + cl];
                b1[ct] = z1[(4 * s + kk) * LDB + ct * 16 + cl];
            }
#pragma unroll
            for (int i = 0; i < RBW; i++)
#pragma unroll
                for (int ct = 0; ct < CT; ct++) {
                    acc0[i][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar0[i][sl], b0[ct], acc0[i][ct], 0, 0, 0);
                    acc1[i][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar1[i][sl], b1[ct], acc1[i][ct], 0, 0, 0);
                }
        }
#pragma unroll
        for (int i = 0; i < RBW; i++)
#pragma unroll
            for (int sl = 0; sl < 4; sl++) {
                ar0[i][sl] = an0[i][sl];
                ar1[i][sl] = an1[i][sl];
            }
    }
}

// Column t < BCOLS of the extra row from the staging threads' partials (mode t/2, part t%2) of
// a block of NW waves, added by row class 0..15 in order (one, two, four waves) or by staging
// thread (three); every thread calls it (it synchronises); uses 2 * 256 doubles of lds.
template <int NW>
__device__ __forceinline__ double reduce_extra(const XAcc<NW>& xacc, double* lds) {
    using XA = XAcc<NW>;
    const int t = threadIdx.x, rt = t >> 4, mt = t & 15;
#pragma unroll
    for (int j = 0; j < XA::N; j++) {
        const int q = rt + XA::RS * ((xacc.base + j) % XA::N);   // the slot's row class
        lds[2 * (q * 16 + mt)] = xacc.s[j].x;
        lds[2 * (q * 16 + mt) + 1] = xacc.s[j].y;
    }
    __syncthreads();
    double v = 0.0;
    if (t < BCOLS) {
        const int ml = t >> 1, part = t & 1;
        constexpr int NQ = XA::BYCLASS ? 16 : XA::RS;
        for (int q = 0; q < NQ; q++) v += lds[2 * (q * 16 + ml) + part];
    }
    __syncthreads();
    return v;
}

// The dense extra row of a back operator is applied by extra_dot from its coefficients parked
// in the padding column BCOLS of the LDS panel it acts on (rows below 4 ks; the contractions
// read only columns below BCOLS).  xrow_fetch reads thread t's coefficient t once per block,
// early (its latency hides behind the forward contraction); xrow_park stores it into the panel
// with the panel's other writes.  Needs 4 ks <= the block's threads and the panel's rows.
__device__ __forceinline__ double xrow_fetch(const FoldOp& op) {
    const int t = threadIdx.x;
    return (op.extra && t < 4 * op.ks) ? gmem(fold_xrow(op))[t] : 0.0;
}
__device__ __forceinline__ void xrow_park(const FoldOp& op, double c, double* z) {
    if (op.extra && int(threadIdx.x) < 4 * op.ks) z[threadIdx.x * LDB + BCOLS] = c;
}
// Column t < BCOLS of the extra row applied to LDS panel z0 (rows 4 ks, stride LDB), its
// coefficients parked in column BCOLS (the sum runs in row order)
__device__ __forceinline__ double extra_dot(const FoldOp& op, const double* z0) {
    const int t = threadIdx.x, n = 4 * op.ks;
    double v = 0.0;
    for (int j = 0; j < n; j++) v += z0[j * LDB + BCOLS] * z0[j * LDB + t];
    return v;
}

// The two row groups of a folded result into LDS panels G0 (rows0 rows) and G1 (rows1 rows),
// row stride LDB; the extra group-0 row (computed elsewhere) is written by the caller.
template <int RBW, int NW, class T = double>
__device__ __forceinline__ void store_groups(const acc_t<T> (&acc0)[RBW][CT], const acc_t<T> (&acc1)[RBW][CT],
                                             const FoldOp& op, double* G0, double* G1) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, cl = lane & 15;
#pragma unroll
    for (int ii = 0; ii < RBW; ii++) {
        const int b = wave + NW * ii;
#pragma unroll
        for (int ct = 0; ct < CT; ct++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = (wave + NW * ii) * 16 + Mfma<T>::row(lane, r);
                if (b < op.rb0 && row < op.rows0) G0[row * LDB + ct * 16 + cl] = acc0[ii][ct][r];
                if (b < op.rb1 && row < op.rows1) G1[row * LDB + ct * 16 + cl] = acc1[ii][ct][r];
            }
    }
}

// Row pair i < n/2 + (n & 1) of a folded result in LDS groups G0, G1 (row stride LDB) at
// column c: (out_i, out_{n-1-i}) = (G0_i + G1_i, G0_i - G1_i); on the centre row of odd n,
// G0_h for both (one node).  Branch free: G1 is read at a clamped row and dropped there.
__device__ __forceinline__ double2 unfold_pair(const double* G0, const double* G1, int n, int i, int c) {
    const int h = n / 2;
    const double g0 = G0[i * LDB + c];
    const double g1r = G1[min(i, h - 1) * LDB + c];
    const double g1 = i < h ? g1r : 0.0;
    return make_double2(g0 + g1, g0 - g1);
}

template <int RBW, class A>
__device__ __forceinline__ void zero_acc2(A (&acc)[RBW][CT]) {
#pragma unroll
    for (int i = 0; i < RBW; i++)
#pragma unroll
        for (int ct = 0; ct < CT; ct++) acc[i][ct] = A{};
}

// Coordinates of acc[i][ct][r] in a block of NW waves: row within the group =
// (wave + NW i)*16 + Mfma<T>::row(lane, r) (the result layout of the instruction),
// column = cb*BCOLS + ct*16 + lane%16.
template <int NW, class T = double>
__device__ __forceinline__ int out_row(int i, int r) {
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    return (wave + NW * i) * 16 + Mfma<T>::row(lane, r);
}
__device__ __forceinline__ int out_col(int ct, int cb) { return cb * BCOLS + ct * 16 + (threadIdx.x & 15); }

// zero rows [rows, zr) of an LDS panel (the k rows past an operator's K that a chunked
// contraction still reads; its A columns there are zero, so the rows must be finite)
template <int NW>
__device__ __forceinline__ void zero_panel_tail(double* z, int rows, int zr) {
    for (int e = threadIdx.x; e < (zr - rows) * BCOLS; e += 64 * NW) z[(rows + e / BCOLS) * LDB + e % BCOLS] = 0.0;
}

// panel rows of the LDS-resident z panels of a back contraction with k-steps ks: 4 rows per
// k-step (fold_gemm_lds reads rows 4 s + 0..3 for s < ks; rows past the operator's K are zeroed)
__host__ __device__ inline int zpanel_rows(int ks) { return 4 * ks; }

}  // namespace dns
