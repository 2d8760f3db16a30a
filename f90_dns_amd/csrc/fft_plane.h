// fft_plane.h -- the LDS-resident plane transforms (device code), included by the
// translation units that instantiate them (fft_c2r64/32.hip, fft_r2c64/32.hip).
//
// 2D plane transforms real(nx,ny) <-> complex(nx/2+1,ny) with the whole half-complex plane
// resident in LDS (one workgroup per plane and field), in FP64 or FP32.
//
// c2r: complex inverse along y for every kx column, then the x direction as a half-length
// complex inverse of packed pairs z_n = x_{2n} + i x_{2n+1}; the packing uses only Re of the
// kx = 0 and kx = nx/2 columns after the y inverse, which is FFTW's c2r result for
// non-Hermitian input (fftw3_dns_module.f90:116-132).  Output divided by nx*ny.
// r2c: the mirror image, unnormalised (fftw3_dns_module.f90:95-108).
//
// Layout and stages.  The plane sits in LDS as M = nx/2 columns x ny rows (row pitch M), the
// y and x twiddle tables after it.  The kx = 0 and kx = nx/2 columns travel as ONE complex
// column: c2r needs only Re of their y inverses, which are the y inverses of their Hermitian
// parts, so column 0 carries herm(X_0) + i herm(X_M); r2c transforms the two real columns
// X_0 + i X_M together and separates them afterwards.  Each direction is a Stockham sequence
// of radix-8/4/3/2 stages (host plan, at least two stages); a stage reads all its inputs to
// registers, then writes its outputs.  The first stage of the first direction reads HBM
// directly, the last stage of the second direction writes HBM directly, and the
// real/half-complex split of the x direction is folded into the loads of the stage next to it.
#pragma once
#include "kernels.h"

namespace dns {

template <int C> struct Cls {
    static constexpr int NT = CLS_NT[C], MAXE = CLS_MAXE[C];
};

template <class T> struct Vec;
template <> struct Vec<double> { using type = double2; };
template <> struct Vec<float> { using type = float2; };
template <class S> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

// x / d for 0 <= x < 2^16, 1 <= d < 2^16 by a multiply-high (ceil(2^32/d) is exact there)
struct FastDiv {
    unsigned m;
    int d;
};
__device__ __forceinline__ FastDiv fastdiv(int d) {
    FastDiv f;
    f.d = d;
    f.m = d > 1 ? 0xFFFFFFFFu / unsigned(d) + 1u : 0u;
    return f;
}
__device__ __forceinline__ int fdiv(int x, const FastDiv& f) {
    return f.d > 1 ? int(__umulhi(unsigned(x), f.m)) : x;
}

// LDS bank swizzle of the flat plane index: x ^ ((x >> 3) & mask) permutes each aligned
// block of 16 (FP64) or 32 (FP32) elements, so that the stride-R writes of a first Stockham stage
// (and the stride-Ns runs of later ones) spread over the banks, while runs of consecutive
// indices stay conflict free.  The plane region is a whole number of such blocks.
template <class V> struct Swz;
template <> struct Swz<double2> { static constexpr int MASK = 15; };
template <> struct Swz<float2> { static constexpr int MASK = 31; };
template <class V> __device__ __forceinline__ int swz(int x) { return x ^ ((x >> 3) & Swz<V>::MASK); }
// complex elements of the plane region (M * ny rounded up to a whole swizzle block)
__host__ __device__ inline int plane_region(int M, int ny) { return (M * ny + 31) & ~31; }

// XCD-aware block order.  Workgroups go to the 8 XCDs round-robin by linear id, so block
// 8 * slot + x runs on XCD x; the slots of one XCD take the nf fields of one plane (or ragged
// tile) one after another.  Blocks that read the same planes (omega_z reads u and v; each cross
// product reads four of the six physical fields) then share an L2 and run close in time.
constexpr int XCDS = 8;
__device__ __forceinline__ bool xcd_item(int n, int nf, int& item, int& f) {
    const int L = blockIdx.x, x = L % XCDS, slot = L / XCDS;
    const int t = slot / nf;
    f = slot - t * nf;
    item = t * XCDS + x;
    return item < n;
}
inline int xcd_grid(int n, int nf) { return XCDS * nf * ((n + XCDS - 1) / XCDS); }

// complex value of type V from a complex element of type S
template <class V, class S>
__device__ __forceinline__ V ldc(const void* p, size_t idx2) {
    using U = typename Scalar<V>::type;
    const typename Vec2<S>::type x = reinterpret_cast<const typename Vec2<S>::type*>(gmem(p))[idx2];
    return mk2(U(x.x), U(x.y));
}
template <class D, class V>
__device__ __forceinline__ void stc(void* p, size_t idx2, V v) {
    reinterpret_cast<typename Vec2<D>::type*>(gmem(p))[idx2] = mk2(D(v.x), D(v.y));
}

// ---- in-register DFTs: out_q = sum_r v_r exp(S 2 pi i q r / R), unnormalised -------------
template <int S, class V> __device__ __forceinline__ V muli(V z) { return S > 0 ? mk2(-z.y, z.x) : mk2(z.y, -z.x); }
// z * (c + S i s)
template <int S, class V> __device__ __forceinline__ V rot(V z, double c, double s) {
    using U = typename Scalar<V>::type;
    const U cc = U(c), ss = U(S > 0 ? s : -s);
    return mk2(z.x * cc - z.y * ss, z.x * ss + z.y * cc);
}
template <int S, class V> __device__ __forceinline__ void dft4(V& a, V& b, V& c, V& d) {
    const V s02 = cadd(a, c), d02 = csub(a, c);
    const V s13 = cadd(b, d), d13 = muli<S>(csub(b, d));
    a = cadd(s02, s13); c = csub(s02, s13);
    b = cadd(d02, d13); d = csub(d02, d13);
}
template <int R, int S, class V> struct PDFT;
template <int S, class V> struct PDFT<2, S, V> {
    static __device__ __forceinline__ void run(V* v) {
        const V a = v[0], b = v[1];
        v[0] = cadd(a, b); v[1] = csub(a, b);
    }
};
template <int S, class V> struct PDFT<3, S, V> {
    static __device__ __forceinline__ void run(V* v) {
        using U = typename Scalar<V>::type;
        const U c = U(-0.5), s = U(0.86602540378443864676372317075293618);   // cos, sin(2 pi/3)
        const V a = v[0], sum = cadd(v[1], v[2]), dif = csub(v[1], v[2]);
        const V t = mk2(a.x + c * sum.x, a.y + c * sum.y);
        const V r = muli<S>(mk2(s * dif.x, s * dif.y));
        v[0] = cadd(a, sum);
        v[1] = cadd(t, r);
        v[2] = csub(t, r);
    }
};
template <int S, class V> struct PDFT<4, S, V> {
    static __device__ __forceinline__ void run(V* v) { dft4<S>(v[0], v[1], v[2], v[3]); }
};
constexpr double RH = 0.70710678118654752440;    // cos(pi/4)
template <int S, class V> struct PDFT<8, S, V> {
    // radix 2 x 4: even and odd halves, then exp(S 2 pi i q/8) on the odd half
    static __device__ __forceinline__ void run(V* v) {
        V e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        V o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4<S>(e0, e1, e2, e3);
        dft4<S>(o0, o1, o2, o3);
        o1 = rot<S>(o1, RH, RH);
        o2 = muli<S>(o2);
        o3 = rot<S>(o3, -RH, RH);
        v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
    }
};
// ---- one Stockham stage ---------------------------------------------------------------------
// nb transforms of length N, radix R, Ns = product of the earlier radices.  Input element e of
// transform b is ld.finish(b, e, ld.fetch(b, e)): every fetch of the stage (the memory reads)
// is issued before any finish (the arithmetic on them), so the reads are in flight together.
// Output element e goes to st(b, e, value).  tw is the LDS
// twiddle table exp(-2 pi i k/L); twstep = L / (Ns R).  BFAST: consecutive threads take
// consecutive transforms (y direction), else consecutive butterflies of one transform.
// SYNC_MID: the loads read LDS that the stores overwrite; SYNC_END: the stores write LDS.
// mid() runs after the fetches are issued and before the first finish (work whose own reads
// then overlap the stage's).
struct NoMid {
    __device__ __forceinline__ void operator()() const {}
};
// the fetch half of a stage: the reads of the thread's MAXB butterflies
template <int R, int NT, int MAXB, bool BFAST, class LD>
__device__ __forceinline__ void stage_fetch(int N, int nb, const LD& ld, typename LD::Raw (&raw)[MAXB][R]) {
    const int m = N / R;
    const int total = nb * m;
    const FastDiv dq = fastdiv(BFAST ? nb : m);
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int t = threadIdx.x + q * NT;
        int b, j;
        if (BFAST) { j = fdiv(t, dq); b = t - j * nb; }
        else { b = fdiv(t, dq); j = t - b * m; }
        if (t < total) {
#pragma unroll
            for (int r = 0; r < R; r++) raw[q][r] = ld.fetch(b, j + r * m);
        }
    }
}
// the rest of a stage from its reads: finish, twiddles, butterflies, stores
template <int R, int NT, int MAXB, int S, bool BFAST, bool SYNC_MID, bool SYNC_END, class V, class LD, class ST>
__device__ __forceinline__ void stage_rest(int N, int Ns, int nb, const V* tw, int twstep, const LD& ld,
                                           const ST& st, const typename LD::Raw (&raw)[MAXB][R]) {
    const int m = N / R;
    const int total = nb * m;
    const FastDiv dq = fastdiv(BFAST ? nb : m), dn = fastdiv(Ns);
    V v[MAXB][R];
    int bi[MAXB], ji[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int t = threadIdx.x + q * NT;
        int b, j;
        if (BFAST) { j = fdiv(t, dq); b = t - j * nb; }
        else { b = fdiv(t, dq); j = t - b * m; }
        bi[q] = b; ji[q] = j;
        if (t < total) {
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = ld.finish(b, j + r * m, raw[q][r]);
        }
    }
    if (SYNC_MID) __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int t = threadIdx.x + q * NT;
        if (t < total) {
            const int b = bi[q], j = ji[q];
            const int kq = fdiv(j, dn), k = j - kq * Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    V w = tw[k * r * twstep];
                    if (S > 0) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            PDFT<R, S, V>::run(v[q]);
            const int base = kq * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) st(b, base + r * Ns, v[q][r]);
        }
    }
    if (SYNC_END) __syncthreads();
}
template <int R, int NT, int MAXB, int S, bool BFAST, bool SYNC_MID, bool SYNC_END, class V, class LD, class ST,
          class MID = NoMid>
__device__ __forceinline__ void plane_stage(int N, int Ns, int nb, const V* tw, int twstep, const LD& ld,
                                            const ST& st, const MID& mid = MID{}) {
    typename LD::Raw raw[MAXB][R];
    stage_fetch<R, NT, MAXB, BFAST>(N, nb, ld, raw);
    mid();
    stage_rest<R, NT, MAXB, S, BFAST, SYNC_MID, SYNC_END>(N, Ns, nb, tw, twstep, ld, st, raw);
}

template <int C, bool F32, int S, bool BFAST, class V, class LD, class ST, class MF = NoMid>
__device__ __forceinline__ void stage_r(int R, int N, int Ns, int nb, const V* tw, int twstep, const LD& ld,
                                        const ST& st, const MF& mid = MF{}) {
    constexpr int NT = Cls<C>::NT;
    constexpr bool MID = LD::kLds && ST::kLds, END = ST::kLds;
    if (R == 8)
        plane_stage<8, NT, stage_maxb(C, 8, F32), S, BFAST, MID, END>(N, Ns, nb, tw, twstep, ld, st, mid);
    else if (R == 4)
        plane_stage<4, NT, stage_maxb(C, 4, F32), S, BFAST, MID, END>(N, Ns, nb, tw, twstep, ld, st, mid);
    else if (R == 3)
        plane_stage<3, NT, stage_maxb(C, 3, F32), S, BFAST, MID, END>(N, Ns, nb, tw, twstep, ld, st, mid);
    else
        plane_stage<2, NT, stage_maxb(C, 2, F32), S, BFAST, MID, END>(N, Ns, nb, tw, twstep, ld, st, mid);
}

// element (b, e) of the LDS plane: y direction (b = column, e = row) or x (b = row, e = column)
template <class V, bool YDIR> struct LdsLoad {
    static constexpr bool kLds = true;
    using Raw = V;
    const V* lds; int pitch;
    __device__ __forceinline__ V fetch(int b, int e) const {
        return lds[swz<V>(YDIR ? e * pitch + b : b * pitch + e)];
    }
    __device__ __forceinline__ V finish(int, int, V x) const { return x; }
};
template <class V, bool YDIR> struct LdsStore {
    static constexpr bool kLds = true;
    V* lds; int pitch;
    __device__ __forceinline__ void operator()(int b, int e, V x) const {
        lds[swz<V>(YDIR ? e * pitch + b : b * pitch + e)] = x;
    }
};

// all stages of one direction (plans have at least two): first-stage loads from ld0, last-stage
// stores to stl, LDS in between; mid as in plane_stage, for the first stage
template <int C, bool F32, int S, bool BFAST, class V, class LD0, class STL, class MF = NoMid>
__device__ __forceinline__ void run_direction(const Plan& P, int nb, V* lds, int pitch, const V* tw, int L,
                                              const LD0& ld0, const STL& stl, const MF& mid = MF{}) {
    const LdsLoad<V, BFAST> ldl{lds, pitch};
    const LdsStore<V, BFAST> stm{lds, pitch};
    int Ns = P.r[0];
    stage_r<C, F32, S, BFAST>(Ns, P.n, 1, nb, tw, L / Ns, ld0, stm, mid);
    for (int s = 1; s < P.ns - 1; s++) {
        const int R = P.r[s];
        stage_r<C, F32, S, BFAST>(R, P.n, Ns, nb, tw, L / (Ns * R), ldl, stm);
        Ns *= R;
    }
    const int R = P.r[P.ns - 1];
    stage_r<C, F32, S, BFAST>(R, P.n, Ns, nb, tw, L / (Ns * R), ldl, stl);
}

// run_direction whose first stage, of radix R0 = P.r[0], already has its reads in raw
template <int C, bool F32, int S, bool BFAST, int R0, class V, class LD0, class STL>
__device__ __forceinline__ void run_direction_pre(const Plan& P, int nb, V* lds, int pitch, const V* tw, int L,
                                                  const LD0& ld0,
                                                  const typename LD0::Raw (&raw)[stage_maxb(C, R0, F32)][R0],
                                                  const STL& stl) {
    const LdsLoad<V, BFAST> ldl{lds, pitch};
    const LdsStore<V, BFAST> stm{lds, pitch};
    stage_rest<R0, Cls<C>::NT, stage_maxb(C, R0, F32), S, BFAST, false, true>(P.n, 1, nb, tw, L / R0, ld0, stm, raw);
    int Ns = R0;
    for (int s = 1; s < P.ns - 1; s++) {
        const int R = P.r[s];
        stage_r<C, F32, S, BFAST>(R, P.n, Ns, nb, tw, L / (Ns * R), ldl, stm);
        Ns *= R;
    }
    const int R = P.r[P.ns - 1];
    stage_r<C, F32, S, BFAST>(R, P.n, Ns, nb, tw, L / (Ns * R), ldl, stl);
}

// twiddle tables of the compute precision into LDS: twy (ny), then twx (nx), contiguous; the
// reads of a batch of four per thread are issued before their LDS writes
template <int NT, class V>
__device__ __forceinline__ void load_twiddles(const FFTArgs& a, V* twy) {
    const V* gy = reinterpret_cast<const V*>(gmem(a.twy));
    const V* gx = reinterpret_cast<const V*>(gmem(a.twx)) - a.ny;
    const int ny = a.ny, n = ny + a.nx;
    for (int base = 0; base < n; base += 4 * NT) {
        // branch free: lanes past the end read and write entry n - 1 again (same value)
        V t[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = min(base + int(threadIdx.x) + u * NT, n - 1);
            t[u] = (e < ny ? gy : gx)[e];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) twy[min(base + int(threadIdx.x) + u * NT, n - 1)] = t[u];
    }
}

// ---- c2r ------------------------------------------------------------------------------------
// What a c2r block transforms: field f as stored (PLAIN), omega_z = i kx v - i ky u from u, v
// (OMZ), or the divergence i kx u + i ky v + D1 w (DIV).  The kind is uniform per block.
enum C2RKind { K_PLAIN = 0, K_OMZ = 1, K_DIV = 2 };

// LDS side buffers of the c2r kernels, after the plane and the twiddle tables: column 0 of the
// combined plane (ny), kx (nx/2+1) and ky (ny) of the compute precision
template <class V> struct C2RSide {
    V* c0;
    typename Scalar<V>::type* kx;
    typename Scalar<V>::type* ky;
};

// Prologue of a c2r block: twiddle tables (load_twiddles layout), column 0 of the combined
// plane, herm(X_0) + i herm(X_M) with herm(x)_y = (x_y + conj x_{-y}) / 2, and the wavenumber
// tables.  Each pass issues all its reads (indices clamped into range, so no lane branches and
// the loads are not sunk into conditional stores), then computes and writes; lanes past the
// end repeat the last entry with its own value.
template <int NT, int KIND, class V, class S>
__device__ __forceinline__ void c2r_prologue(const FFTArgs& a, int f, size_t p2, V* twy, const C2RSide<V>& sd) {
    using U = typename Scalar<V>::type;
    const int ny = a.ny, M = a.M, ntw = ny + a.nx;
    const V* gy = reinterpret_cast<const V*>(gmem(a.twy));
    const V* gx = reinterpret_cast<const V*>(gmem(a.twx)) - ny;
    const int src0 = KIND == K_PLAIN ? f : 0;
    const int passes = max(max((ntw + 4 * NT - 1) / (4 * NT), (ny + NT - 1) / NT), (M + NT) / NT);
    for (int p = 0; p < passes; p++) {
        const int t0 = p * 4 * NT + int(threadIdx.x), y = min(p * NT + int(threadIdx.x), ny - 1);
        const int i = min(p * NT + int(threadIdx.x), M);
        V t[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = min(t0 + u * NT, ntw - 1);
            t[u] = (e < ny ? gy : gx)[e];
        }
        const int yr = y == 0 ? 0 : ny - y;
        const size_t o[4] = {p2 + size_t(y) * a.nxh, p2 + size_t(yr) * a.nxh, p2 + size_t(y) * a.nxh + M,
                             p2 + size_t(yr) * a.nxh + M};
        V uu[4], vv[4], dw[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uu[q] = ldc<V, S>(a.src[src0], o[q]);
            if (KIND != K_PLAIN) vv[q] = ldc<V, S>(a.src[1], o[q]);
            if (KIND == K_DIV) dw[q] = ldc<V, S>(a.src[2], o[q]);
        }
        // wavenumbers only for OMZ / DIV (a plain transform's xw, yw may be null)
        U k0 = U(0), km = U(0), ky = U(0), kyr = U(0), kxi = U(0);
        if constexpr (KIND != K_PLAIN) {
            const double* xw = gmem(a.xw);
            const double* yw = gmem(a.yw);
            k0 = U(xw[0]); km = U(xw[M]); ky = U(yw[y]); kyr = U(yw[yr]); kxi = U(xw[i]);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) twy[min(t0 + u * NT, ntw - 1)] = t[u];
        V x[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const U kx = q < 2 ? k0 : km, kyq = (q & 1) ? kyr : ky;
            if (KIND == K_PLAIN) x[q] = uu[q];
            else if (KIND == K_OMZ) x[q] = mk2(-kx * vv[q].y - (-kyq * uu[q].y), kx * vv[q].x - kyq * uu[q].x);
            else x[q] = mk2(-kx * uu[q].y + -kyq * vv[q].y + dw[q].x, kx * uu[q].x + kyq * vv[q].x + dw[q].y);
        }
        const U h = U(0.5);
        const V h0 = mk2(h * (x[0].x + x[1].x), h * (x[0].y - x[1].y));
        const V hm = mk2(h * (x[2].x + x[3].x), h * (x[2].y - x[3].y));
        sd.c0[y] = mk2(h0.x - hm.y, h0.y + hm.x);
        if constexpr (KIND != K_PLAIN) {
            sd.ky[y] = ky;
            sd.kx[i] = kxi;
        }
    }
}

// first y stage of c2r: column b, row y of the combined plane; columns b > 0 from HBM (the
// reads are issued by fetch, the arithmetic of OMZ / DIV is done by finish), column 0 from the
// side buffer
template <int KIND, class V> struct C2RRaw { V u; };   // converted to the compute type on reading
template <class V> struct C2RRaw<K_OMZ, V> { V u, v; };
template <class V> struct C2RRaw<K_DIV, V> { V u, v, dw; };

template <int KIND, class V, class S> struct C2RLoad {
    static constexpr bool kLds = false;
    using Raw = C2RRaw<KIND, V>;
    const FFTArgs* a; int f; size_t p2; C2RSide<V> sd;
    // branch free: every lane reads its HBM element (column 0 lanes read X_0, unused); finish
    // takes column 0 from the side buffer (the lanes of a row read one address)
    __device__ __forceinline__ Raw fetch(int b, int y) const {
        Raw r;
        const size_t off2 = p2 + size_t(y) * a->nxh + b;
        if constexpr (KIND == K_PLAIN) {
            r.u = ldc<V, S>(a->src[f], off2);
        } else {
            r.u = ldc<V, S>(a->src[0], off2);
            r.v = ldc<V, S>(a->src[1], off2);
            if constexpr (KIND == K_DIV) r.dw = ldc<V, S>(a->src[2], off2);
        }
        return r;
    }
    __device__ __forceinline__ V finish(int b, int y, const Raw& r) const {
        V z;
        if constexpr (KIND == K_PLAIN) {
            z = r.u;
        } else {
            using U = typename Scalar<V>::type;
            const U kx = sd.kx[b], ky = sd.ky[y];
            const V u = r.u, v = r.v;
            if constexpr (KIND == K_OMZ) z = mk2(-kx * v.y - (-ky * u.y), kx * v.x - ky * u.x);
            else z = mk2(-kx * u.y + -ky * v.y + r.dw.x, kx * u.x + ky * v.x + r.dw.y);
        }
        return b > 0 ? z : sd.c0[y];
    }
};

// first x stage of c2r: packed z_n of row y from the y-transformed plane
template <class V> struct PackLoad {
    static constexpr bool kLds = true;
    using Raw = V;
    const V* lds; const V* twx; int M;
    __device__ __forceinline__ V finish(int, int, V x) const { return x; }
    __device__ __forceinline__ V fetch(int y, int n) const {
        const int r0 = y * M;
        const V A = lds[swz<V>(r0 + n)], B = lds[swz<V>(r0 + (n == 0 ? 0 : M - n))];
        const V tw = twx[n];
        const V w = mk2(tw.x, -tw.y);                    // exp(+2 pi i n/nx)
        const V sm = mk2(A.x + B.x, A.y - B.y);          // A + conj B
        const V d = mk2(A.x - B.x, A.y + B.y);           // A - conj B
        const V wd = cmul(w, d);
        const V z = mk2(sm.x - wd.y, sm.y + wd.x);       // s + i w d
        // n = 0: column 0 holds (Re Y_0, Re Y_M): z_0 = (Y_0 + Y_M, Y_0 - Y_M)
        return n == 0 ? mk2(A.x + A.y, A.x - A.y) : z;
    }
};

// last x stage of c2r: z_n / (nx ny) to the real row y
template <class V, class D> struct C2RStore {
    static constexpr bool kLds = false;
    void* out; size_t ob2; int M; typename Scalar<V>::type inv;
    __device__ __forceinline__ void operator()(int y, int n, V z) const {
        stc<D>(out, ob2 + size_t(y) * M + n, mk2(z.x * inv, z.y * inv));
    }
};

template <int C, int KIND, class T, class SRC, class DST>
__device__ __forceinline__ void c2r_body(const FFTArgs& a, int k, int f, typename Vec<T>::type* lds) {
    using V = typename Vec<T>::type;
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int NT = Cls<C>::NT;
    const int M = a.M, ny = a.ny, nx = a.nx;
    V* twy = lds + plane_region(M, ny);
    V* twx = twy + ny;
    V* c0 = twx + nx;
    T* kx = reinterpret_cast<T*>(c0 + ny);
    const C2RSide<V> sd{c0, kx, kx + a.nxh};
    const size_t p2 = (size_t(k) * a.sstride) / 2;
    // A plain transform runs the prologue inside the first y stage, after the stage's reads are
    // issued, so that both sets of reads are in flight together (the first stage uses no
    // twiddles; its finish takes column 0 from the side buffer the prologue writes).  omega_z
    // and the divergence hold two or three reads per element: their prologue runs first (the
    // held reads would spill).
    auto prologue = [&]() {
        c2r_prologue<NT, KIND, V, SRC>(a, f, p2, twy, sd);
        __syncthreads();
    };
    const C2RLoad<KIND, V, SRC> ld0{&a, f, p2, sd};
    if constexpr (KIND == K_PLAIN) {
        run_direction<C, F32, +1, true>(a.py, M, lds, M, twy, ny, ld0, LdsStore<V, true>{lds, M}, prologue);
    } else {
        prologue();
        run_direction<C, F32, +1, true>(a.py, M, lds, M, twy, ny, ld0, LdsStore<V, true>{lds, M});
    }
    const T inv = T(1) / (T(nx) * T(ny));
    run_direction<C, F32, +1, false>(a.px, ny, lds, M, twx, nx, PackLoad<V>{lds, twx, M},
                                     C2RStore<V, DST>{a.dst[f], (size_t(k) * a.dstride) / 2, M, inv});
}

// plane k of field f.  Storage types: FP64 compute reads and writes FP64; FP32 compute of the
// step (NL) reads u, v, w in FP64 and omega_x, omega_y in FP32 and writes FP32; FP32 PLAIN
// (the FFT module) reads and writes FP64.
template <int C, int OP, class T>
__device__ __forceinline__ void c2r_plane(const FFTArgs& a, int k, int f) {
    using V = typename Vec<T>::type;
    extern __shared__ char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    if constexpr (OP == C2R_PLAIN) {
        c2r_body<C, K_PLAIN, T, double, double>(a, k, f, lds);
    } else if constexpr (OP == C2R_DIV) {
        c2r_body<C, K_DIV, T, double, double>(a, k, f, lds);
    } else if constexpr (sizeof(T) == 8) {
        if (f < 5) c2r_body<C, K_PLAIN, T, double, double>(a, k, f, lds);
        else c2r_body<C, K_OMZ, T, double, double>(a, k, f, lds);
    } else {
        // omega_z reads u, v (bit 0)
        if (f == 5) c2r_body<C, K_OMZ, T, double, float>(a, k, f, lds);
        else if ((a.src_f32 >> f) & 1) c2r_body<C, K_PLAIN, T, float, float>(a, k, f, lds);
        else c2r_body<C, K_PLAIN, T, double, float>(a, k, f, lds);
    }
}

// plane of item k of a direct launch (FFTArgs::zin)
__device__ __forceinline__ int item_plane(const FFTArgs& a, int k) {
    if (a.zin == 0) return k;
    const int m = k / a.zin;
    return m * (a.zin + 2) + 1 + (k - m * a.zin);
}

// grid of xcd_grid(np, nf) blocks over np items (planes, or ragged tiles) and nf fields
template <int C, int OP, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_c2r(FFTArgs a, int np, int nf) {
    int k, f;
    if (xcd_item(np, nf, k, f)) c2r_plane<C, OP, T>(a, item_plane(a, k), f);
}
template <int C, int OP, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_c2r_rag(const FFTArgs* __restrict__ d, const Tile* __restrict__ map,
                                                        int nt, int nf) {
    int i, f;
    if (!xcd_item(nt, nf, i, f)) return;
    const Tile t = map[i];
    const FFTArgs& a = d[t.g];
    c2r_plane<C, OP, T>(a, t.x, f);
}

// ---- c2r of the step, persistent ---------------------------------------------------------------
// A block per CU walks the step's c2r items (plane, field) in the XCD-aware order of the direct
// launch: virtual block L = blockIdx.x, blockIdx.x + gridDim.x, ... (gridDim.x a multiple of 8,
// so item L still runs on XCD L % 8).  The twiddle and wavenumber tables are loaded once.  Once
// an item's y direction is done, the reads of the next item's first stage (when it is a plain
// field) are issued into registers, so they travel while this item's x direction runs in LDS
// and its real rows are written.  (Issued right after the first y stage they stalled the y
// stages: the kernel spills a few registers, and a spill reload waits for every read issued
// before it.)

// wavenumber tables of the side buffer (kx: nx/2 + 1, ky: ny) of the compute precision
template <int NT, class V>
__device__ __forceinline__ void c2r_wavenumbers(const FFTArgs& a, const C2RSide<V>& sd) {
    using U = typename Scalar<V>::type;
    const double* xw = gmem(a.xw);
    const double* yw = gmem(a.yw);
    for (int i = threadIdx.x; i < a.nxh; i += NT) sd.kx[i] = U(xw[i]);
    for (int y = threadIdx.x; y < a.ny; y += NT) sd.ky[y] = U(yw[y]);
}

// column 0 of the combined plane, herm(X_0) + i herm(X_M), into the side buffer (the item part
// of c2r_prologue; the wavenumbers come from the side buffer)
template <int NT, int KIND, class V, class S>
__device__ __forceinline__ void c2r_column0(const FFTArgs& a, int f, size_t p2, const C2RSide<V>& sd) {
    using U = typename Scalar<V>::type;
    const int ny = a.ny, M = a.M;
    const int src0 = KIND == K_PLAIN ? f : 0;
    for (int y0 = 0; y0 < ny; y0 += NT) {
        const int y = min(y0 + int(threadIdx.x), ny - 1), yr = y == 0 ? 0 : ny - y;
        const size_t o[4] = {p2 + size_t(y) * a.nxh, p2 + size_t(yr) * a.nxh, p2 + size_t(y) * a.nxh + M,
                             p2 + size_t(yr) * a.nxh + M};
        V uu[4], vv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uu[q] = ldc<V, S>(a.src[src0], o[q]);
            if (KIND == K_OMZ) vv[q] = ldc<V, S>(a.src[1], o[q]);
        }
        V x[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if constexpr (KIND == K_PLAIN) {
                x[q] = uu[q];
            } else {
                const U kx = q < 2 ? sd.kx[0] : sd.kx[M], kyq = (q & 1) ? sd.ky[yr] : sd.ky[y];
                x[q] = mk2(-kx * vv[q].y - (-kyq * uu[q].y), kx * vv[q].x - kyq * uu[q].x);
            }
        }
        const U h = U(0.5);
        const V h0 = mk2(h * (x[0].x + x[1].x), h * (x[0].y - x[1].y));
        const V hm = mk2(h * (x[2].x + x[3].x), h * (x[2].y - x[3].y));
        sd.c0[y] = mk2(h0.x - hm.y, h0.y + hm.x);
    }
}

// virtual block L of xcd_grid(n, nf): its item and field (xcd_item for a given L)
__device__ __forceinline__ bool xcd_decode(int L, int n, int nf, int& item, int& f) {
    const int x = L % XCDS, slot = L / XCDS;
    const int t = slot / nf;
    f = slot - t * nf;
    item = t * XCDS + x;
    return item < n;
}

// C2R_NL (u, v, w, omega_x, omega_y plain, omega_z from u, v) over np items, first y stage of
// radix R0; nv = xcd_grid(np, 6) virtual blocks.  T = double: FP64 in and out; T = float: the
// storage of the mixed step (u, v, w FP64, omega_x, omega_y as src_f32 marks them, FP32 out).
template <int C, int R0, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_c2r_nl_pers(FFTArgs a, int np, int nv) {
    using V = typename Vec<T>::type;
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int NT = Cls<C>::NT, MB0 = stage_maxb(C, R0, F32), NF = 6;
    extern __shared__ char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    const int M = a.M, ny = a.ny, nx = a.nx;
    V* twy = lds + plane_region(M, ny);
    V* twx = twy + ny;
    V* c0 = twx + nx;
    T* kx = reinterpret_cast<T*>(c0 + ny);
    const C2RSide<V> sd{c0, kx, kx + a.nxh};
    load_twiddles<NT>(a, twy);
    c2r_wavenumbers<NT>(a, sd);
    __syncthreads();
    using LdD = C2RLoad<K_PLAIN, V, double>;
    using LdF = C2RLoad<K_PLAIN, V, float>;
    C2RRaw<K_PLAIN, V> raw[MB0][R0];
    bool have = false;   // raw holds the reads of this item's first stage
    const T inv = T(1) / (T(nx) * T(ny));
    // plain field f stored as float (only with FP32 compute)
    auto src_float = [&](int f) { return F32 && ((a.src_f32 >> f) & 1u) != 0; };
    for (int L = blockIdx.x; L < nv; L += gridDim.x) {
        int k, f;
        if (!xcd_decode(L, np, NF, k, f)) continue;
        const int kp = item_plane(a, k);
        const size_t p2 = (size_t(kp) * a.sstride) / 2;
        if (f < 5) {
            const LdD ldd{&a, f, p2, sd};
            const LdF ldf{&a, f, p2, sd};
            const bool sf = src_float(f);
            if (!have) {
                if (sf) stage_fetch<R0, NT, MB0, true>(a.py.n, M, ldf, raw);
                else stage_fetch<R0, NT, MB0, true>(a.py.n, M, ldd, raw);
            }
            if (sf) c2r_column0<NT, K_PLAIN, V, float>(a, f, p2, sd);
            else c2r_column0<NT, K_PLAIN, V, double>(a, f, p2, sd);
            __syncthreads();
            // finish of a plain load converts nothing (the read did): either loader will do
            run_direction_pre<C, F32, +1, true, R0>(a.py, M, lds, M, twy, ny, ldd, raw, LdsStore<V, true>{lds, M});
        } else {
            c2r_column0<NT, K_OMZ, V, double>(a, f, p2, sd);
            __syncthreads();
            run_direction<C, F32, +1, true>(a.py, M, lds, M, twy, ny, C2RLoad<K_OMZ, V, double>{&a, f, p2, sd},
                                            LdsStore<V, true>{lds, M});
        }
        // the next item's first-stage reads, when it is a plain field
        {
            int kn, fn;
            have = xcd_decode(L + gridDim.x, np, NF, kn, fn) && fn < 5;
            if (have) {
                const size_t pn = (size_t(item_plane(a, kn)) * a.sstride) / 2;
                if (src_float(fn)) stage_fetch<R0, NT, MB0, true>(a.py.n, M, LdF{&a, fn, pn, sd}, raw);
                else stage_fetch<R0, NT, MB0, true>(a.py.n, M, LdD{&a, fn, pn, sd}, raw);
            }
        }
        run_direction<C, F32, +1, false>(a.px, ny, lds, M, twx, nx, PackLoad<V>{lds, twx, M},
                                         C2RStore<V, T>{a.dst[f], (size_t(kp) * a.dstride) / 2, M, inv});
        __syncthreads();   // the plane and column 0 are rewritten by the next item
    }
}

// ---- r2c ------------------------------------------------------------------------------------
// first x stage of r2c: packed pair n of row y from HBM (the rotational product for CROSS)
template <int OP, class V, class S> struct R2CLoad {
    static constexpr bool kLds = false;
    using Raw = V;
    const FFTArgs* a; int f; size_t p2;
    __device__ __forceinline__ V finish(int, int, V x) const { return x; }
    __device__ __forceinline__ V fetch(int y, int n) const {
        const size_t q2 = p2 + size_t(y) * a->M + n;
        if (OP == R2C_PLAIN) return ldc<V, S>(a->src[f], q2);
        // rotational product S = u x omega (DNS_pressure_BC_3D.f90:1089-1091);
        // src = u, v, w, omega_x, omega_y, omega_z
        int ia, ib, ic, id;
        if (f == 0) { ia = 1; ib = 5; ic = 2; id = 4; }        // v*oz - w*oy
        else if (f == 1) { ia = 2; ib = 3; ic = 0; id = 5; }   // w*ox - u*oz
        else { ia = 0; ib = 4; ic = 1; id = 3; }               // u*oy - v*ox
        const V A = ldc<V, S>(a->src[ia], q2);
        const V B = ldc<V, S>(a->src[ib], q2);
        const V Cc = ldc<V, S>(a->src[ic], q2);
        const V D = ldc<V, S>(a->src[id], q2);
        return mk2(A.x * B.x - Cc.x * D.x, A.y * B.y - Cc.y * D.y);
    }
};

// first y stage of r2c: the half-complex X_b of row y from the packed row transforms;
// column 0 gives X_0 + i X_M (both real)
template <class V> struct SplitLoad {
    static constexpr bool kLds = true;
    using Raw = V;
    const V* lds; const V* twx; int M;
    __device__ __forceinline__ V finish(int, int, V x) const { return x; }
    __device__ __forceinline__ V fetch(int b, int y) const {
        using U = typename Scalar<V>::type;
        const int r0 = y * M;
        const V A = lds[swz<V>(r0 + b)], B = lds[swz<V>(r0 + (b == 0 ? 0 : M - b))];
        const V w = twx[b];                                  // exp(-2 pi i b/nx)
        const V sm = mk2(A.x + B.x, A.y - B.y);
        const V d = mk2(A.x - B.x, A.y + B.y);
        const V wd = cmul(w, d);
        const U h = U(0.5);
        const V x = mk2(h * (sm.x + wd.y), h * (sm.y - wd.x));   // (s - i w d)/2
        return b == 0 ? mk2(A.x + A.y, A.x - A.y) : x;
    }
};

// last y stage of r2c: columns b > 0 to HBM (dealias rule applied), column 0 back to LDS
template <class V, class D> struct R2CStore {
    static constexpr bool kLds = true;
    V* lds; void* out; size_t ob2; Dealias dl; int M, nxh;
    __device__ __forceinline__ void operator()(int b, int y, V x) const {
        using U = typename Scalar<V>::type;
        if (b == 0) {
            lds[swz<V>(y * M)] = x;
        } else {
            const size_t e = size_t(y) * nxh + b;
            if (!dl.keep(b, y)) x = mk2(U(0), U(0));
            stc<D>(out, ob2 + e, x);
        }
    }
};

template <int C, int OP, class T, class SRC, class DST>
__device__ __forceinline__ void r2c_body(const FFTArgs& a, int k, int f, typename Vec<T>::type* lds) {
    using V = typename Vec<T>::type;
    using U = T;
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int NT = Cls<C>::NT;
    const int M = a.M, ny = a.ny, nxh = a.nxh;
    V* twy = lds + plane_region(M, ny);
    V* twx = twy + ny;
    run_direction<C, F32, -1, false>(a.px, ny, lds, M, twx, a.nx,
                                     R2CLoad<OP, V, SRC>{&a, f, (size_t(k) * a.sstride) / 2},
                                     LdsStore<V, false>{lds, M});
    const size_t ob2 = (size_t(k) * a.dstride) / 2;
    void* out = a.dst[f];
    const Dealias dl{a.mask != nullptr, a.kxm, a.kym, ny};
    run_direction<C, F32, -1, true>(a.py, M, lds, M, twy, ny, SplitLoad<V>{lds, twx, M},
                                    R2CStore<V, DST>{lds, out, ob2, dl, M, nxh});
    // separate column 0: C_y = F0_y + i FM_y with F0, FM Hermitian in y
    for (int y = threadIdx.x; y < ny; y += NT) {
        const V c = lds[swz<V>(y * M)], cr = lds[swz<V>((y == 0 ? 0 : ny - y) * M)];
        const U h = U(0.5);
        V x0 = mk2(h * (c.x + cr.x), h * (c.y - cr.y));
        V xm = mk2(h * (c.y + cr.y), -h * (c.x - cr.x));
        const size_t e = size_t(y) * nxh;
        if (!dl.keep(0, y)) x0 = mk2(U(0), U(0));
        if (!dl.keep(M, y)) xm = mk2(U(0), U(0));
        stc<DST>(out, ob2 + e, x0);
        stc<DST>(out, ob2 + e + M, xm);
    }
}

// storage types: FP64 compute FP64 in and out; FP32 CROSS (the step) FP32 in and out; FP32
// PLAIN (the FFT module) FP64 in and out
template <int C, int OP, class T>
__device__ __forceinline__ void r2c_plane(const FFTArgs& a, int k, int f) {
    using V = typename Vec<T>::type;
    extern __shared__ char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);
    load_twiddles<Cls<C>::NT>(a, lds + plane_region(a.M, a.ny));
    __syncthreads();
    if constexpr (sizeof(T) == 4 && OP == R2C_CROSS) r2c_body<C, OP, T, float, float>(a, k, f, lds);
    else r2c_body<C, OP, T, double, double>(a, k, f, lds);
}

template <int C, int OP, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_r2c(FFTArgs a, int np, int nf) {
    int k, f;
    if (xcd_item(np, nf, k, f)) r2c_plane<C, OP, T>(a, item_plane(a, k), f);
}
template <int C, int OP, class T>
__global__ __launch_bounds__(Cls<C>::NT) void k_r2c_rag(const FFTArgs* __restrict__ d, const Tile* __restrict__ map,
                                                        int nt, int nf) {
    int i, f;
    if (!xcd_item(nt, nf, i, f)) return;
    const Tile t = map[i];
    const FFTArgs& a = d[t.g];
    r2c_plane<C, OP, T>(a, t.x, f);
}

// ---- launchers of the instantiating translation units ----------------------------------------
// cls = size class; lds = dynamic LDS bytes (plane_lds_bytes); grid (np planes, nf fields)
hipError_t c2r_launch64(int cls, int op, const FFTArgs& a, int np, int nf, size_t lds, hipStream_t s);
hipError_t c2r_launch32(int cls, int op, const FFTArgs& a, int np, int nf, size_t lds, hipStream_t s);
hipError_t r2c_launch64(int cls, int op, const FFTArgs& a, int np, int nf, size_t lds, hipStream_t s);
hipError_t r2c_launch32(int cls, int op, const FFTArgs& a, int np, int nf, size_t lds, hipStream_t s);
hipError_t c2r_launch_rag64(int cls, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds, hipStream_t s);
hipError_t c2r_launch_rag32(int cls, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds, hipStream_t s);
hipError_t r2c_launch_rag64(int cls, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds, hipStream_t s);
hipError_t r2c_launch_rag32(int cls, const FFTArgs* d, const Tile* map, int nt, int nf, size_t lds, hipStream_t s);
// raise the dynamic-LDS limit of every kernel of the unit (once per process)
hipError_t c2r_prepare64();
hipError_t c2r_prepare32();
hipError_t r2c_prepare64();
hipError_t r2c_prepare32();

}  // namespace dns
