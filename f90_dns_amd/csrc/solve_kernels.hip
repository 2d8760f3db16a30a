// solve_kernels.hip -- wall-normal (z) work on spectral columns: the LGL D1 contraction,
// the Crank-Nicolson Helmholtz solves and the pressure Poisson solve + projection, each as
// FP64 MFMA contractions over 16-column tiles per wavefront, plus the small kernels of the
// step (mode (0,0) Poisson, correction, wall values, PI forcing, reductions).
//
// Spectral fields are complex(nx/2+1, ny, nz), plane-major; seen as real matrices they are
// nz rows x H2 = 2 (nx/2+1) ny columns, column c = 2*mode + part (part 0 = Re, 1 = Im).
#include <algorithm>
#include <type_traits>

#include "contract.h"
#include "kernels.h"

namespace dns {

static_assert(BMODES == 16, "staging assigns 16 modes x 16 rows to a block's threads");
static_assert(KC == 16, "a chunk is 16 panel rows");

// waves per SIMD the contraction kernels are compiled for: three with one row-block pair per
// wave (at most 168 VGPRs; four spilled and measured no faster), two with two pairs per wave
#ifndef ZW_D1
#define ZW_D1 3
#endif
#ifndef ZW_HELM
#define ZW_HELM 3
#endif
#ifndef ZW_POIS
#define ZW_POIS 3
#endif
#define Z_WAVES(rbw, w) ((rbw) == 1 ? (w) : 2)
// row pairs whose global reads a thread of the one- to three-wave epilogues issues together
#ifndef ZD_PP
#define ZD_PP 1
#endif
#ifndef POIS_PP
#define POIS_PP 1
#endif

// Wavenumber seen by a derivative whose result is projected by a c2r/r2c round trip:
// the x Nyquist column gives an anti-Hermitian column (zero after the round trip), and so
// does the y Nyquist row in the kx = 0 and kx = nx/2 columns.
__device__ __forceinline__ double kx_eff(const double* xw, int i, int nx) { return i == nx / 2 ? 0.0 : xw[i]; }
__device__ __forceinline__ double ky_eff(const double* yw, int i, int j, int nx, int ny) {
    return (j == ny / 2 && (i == 0 || i == nx / 2)) ? 0.0 : yw[j];
}
static_assert(BMODES == Z_BLOCK_MODES, "column blocks of the tile mode space");

// storage of a tile column kept for its stores: x = so, y = s2o (zcol_t)
__device__ __forceinline__ int2 zst(const ZCol& z) { return make_int2(z.so, z.s2o); }
// store v at the storage of a tile column (at row offset ro), and at its second column: the
// mirrored mode's part (+v real, -v imaginary) or the zeroed imaginary part (s2o = so + 1)
// (sec: the block holds such columns, block-uniform)
__device__ __forceinline__ void zstore(double* out, size_t ro, int2 z, double v, bool sec) {
    out[ro + z.x] = v;
    if (sec && z.y >= 0) out[ro + z.y] = z.y == z.x + 1 ? 0.0 : ((z.x & 1) ? -v : v);
}
__device__ __forceinline__ void zstore_f(float* out, size_t ro, int2 z, double v, bool sec) {
    out[ro + z.x] = float(v);
    if (sec && z.y >= 0) out[ro + z.y] = z.y == z.x + 1 ? 0.0f : float((z.x & 1) ? -v : v);
}
// the same at element q = ro + z.x
__device__ __forceinline__ void zstore_q(double* out, size_t q, int2 z, double v, bool sec) {
    zstore(out, q - z.x, z, v, sec);
}
// the flags of column block cb (ZB_SECOND, ZB_PAIRED), as a wave-uniform value
__device__ __forceinline__ int zflags(const ZMap& z, int cb) { return __builtin_amdgcn_readfirstlane(gmem(z.bfl)[cb]); }

// -------------------------------------------------------------------------------------
// z derivative: out = D1 in  (compute_derivatives_3d dz, DNS_pressure_BC_3D.f90:2348-2358),
// or the vorticity components that need it: omega_x = i ky w - D1 v, omega_y = D1 u - i kx w
// (DNS_pressure_BC_3D.f90:1077-1079).  Folded: s = D1 x_odd on X_O (rows up to the centre),
// r = D1 x_even on X_E, out_k = s_k + r_k, out_{nz-1-k} = s_k - r_k.  One block = 16 modes.
// -------------------------------------------------------------------------------------
// PS: the block holds a paired slot (its two parts are the real parts of two modes).
// T: the contraction's element type (float: FP32 panels, operator and MFMA, FP32 accumulation).
template <int RBW, int MODE, int NW, bool PS, class T = double>
__device__ __forceinline__ void zderiv_tile(const ZArgs& a, int cb, int by, int mb, double* lds) {
    constexpr bool LX = NW < 4;   // fewer row-block pairs than four: one wave per pair, LDS unfold
    constexpr int NT = 64 * NW;
    const int H2 = a.H2, nz = a.nz, hF = nz / 2, LE = hF + (nz & 1);
    const int wave = threadIdx.x >> 6;
    // mode space: the per-mode solves' slots for the vorticity of the step (packed), else the
    // storage modes
    const int P = a.packed ? a.zm.P : H2 / 2, NC = 2 * P;
    // (the one- to three-wave variants are compiled for the storage columns only)
    const int bfl = (NW == 4 && a.packed) ? zflags(a.zm, cb) : ZB_IDENT;
    const bool sec = (bfl & ZB_SECOND) != 0, ident = (bfl & ZB_IDENT) != 0;
    auto zc_of = [&](int col) -> ZCol {
        if (ident) {
            const int q = col >> 1, j = q / a.nxh;
            return ZCol{col, -1, q - j * a.nxh, j};
        }
        return zcol_t(a.zm, col, false);
    };
    const int mt = threadIdx.x & 15, slot = cb * BMODES + mt;
    const bool sval = slot < P;
    const ZCol zx = zc_of(2 * min(slot, P - 1)), zy = zc_of(2 * min(slot, P - 1) + 1);
    const int smode = sval ? (zx.so >> 1) : 0, ox = sval ? zx.so : 0, oy = sval ? zy.so : 0;
    const size_t rs = size_t(H2) / 2;
    // vorticity: one component per block, by = 0: omega_x from v, 1: omega_y from u (two blocks
    // per column block fill the CUs: one block doing both left a third of the slots empty).  Its
    // i k w terms take the effective wavenumbers: they differ from the raw ones only where the
    // term is anti-Hermitian, which the c2r that reads omega drops.
    const size_t mo = size_t(mb) * a.mstride;
    const double* w_in = gmem(a.in[2]) + mo;
    {
        const int f = MODE == Z_VORT ? (by == 0 ? 1 : 0) : by;
        const double* x1 = gmem(a.in[f]) + mo;
        const double2* x2 = reinterpret_cast<const double2*>(x1) + smode;
        // branch-free: mirrored rows (clamped in range), the centre row has na == nb
        struct Raw { double2 xa, xb; };
        auto fetch = [&](int kf) -> Raw {
            const int na = min(kf, nz - 1), nb = max(nz - 1 - kf, 0);
            if constexpr (PS) {
                return Raw{make_double2(x1[size_t(na) * H2 + ox], x1[size_t(na) * H2 + oy]),
                           make_double2(x1[size_t(nb) * H2 + ox], x1[size_t(nb) * H2 + oy])};
            } else {
                return Raw{x2[na * rs], x2[nb * rs]};
            }
        };
        auto finish = [&](int kf, const Raw& rw) -> FoldPair {
            const double2 xa = rw.xa, xb = rw.xb;
            const double keep = (sval && kf < LE) ? 1.0 : 0.0;
            const double pair = kf < hF ? 1.0 : 0.0;   // 0 on the centre row
            FoldPair r;
            r.p0 = make_double2(keep * pair * (xa.x - xb.x), keep * pair * (xa.y - xb.y));
            r.p1 = make_double2(keep * (xa.x + pair * xb.x), keep * (xa.y + pair * xb.y));
            return r;
        };
        const int o = f == 1 ? 0 : 1;
        // the vorticity from y = D1 x at row offset ro of a column: w's two parts of the mode
        // (wc this column's, wo the other), the column's effective wavenumber kk and part
        auto vort = [&](double y, double wc, double wo, double kk, int part) {
            const double wr = part == 0 ? wc : wo, wi = part == 0 ? wo : wc;
            if (f == 1) return part == 0 ? (-kk * wi - y) : (kk * wr - y);   // omega_x = i ky w - D1 v
            return part == 0 ? (y + kk * wi) : (y - kk * wr);                // omega_y = D1 u - i kx w
        };
        auto kk_of = [&](const ZCol& z) {
            return f == 1 ? ky_eff(gmem(a.yw), z.i, z.j, a.nx, a.ny) : kx_eff(gmem(a.xw), z.i, a.nx);
        };
        auto put = [&](size_t ro, int2 z, double v) {
            if (MODE == Z_PLAIN) zstore(gmem(a.out[f]) + mo, ro, z, v, sec);
            else if (a.out_f32) zstore_f(reinterpret_cast<float*>(gmem(a.out[o])) + mo, ro, z, v, sec);
            else zstore(gmem(a.out[o]) + mo, ro, z, v, sec);
        };
        acc_t<T> acc0[RBW][CT], acc1[RBW][CT];
        zero_acc2(acc0);
        zero_acc2(acc1);
        XAcc<NW> xacc;
        fold_gemm<RBW, false, NW, T>(acc0, acc1, a.d1, fetch, finish, reinterpret_cast<T*>(lds), xacc);
        const double xs = a.d1.extra ? reduce_extra<NW>(xacc, lds) : 0.0;   // the centre row s_h
        if constexpr (LX) {
            // few row blocks: unfold through LDS so that every thread stores
            double* G0 = lds;
            double* G1 = lds + a.d1.rows0 * LDB;
            store_groups<RBW, NW, T>(acc0, acc1, a.d1, G0, G1);
            if (a.d1.extra && threadIdx.x < BCOLS) G0[(a.d1.rows0 - 1) * LDB + threadIdx.x] = xs;
            __syncthreads();
            // each thread keeps one column (NT is a multiple of BCOLS) and walks row pairs
            // i = t / BCOLS, + NT / BCOLS, ...: nodes i and nz-1-i from one pair of LDS reads; the
            // vorticity's w reads for PP row pairs go out first (rows clamped into range)
            constexpr int NPR = NT / BCOLS, PP = ZD_PP;
            const int c = threadIdx.x % BCOLS, col = cb * BCOLS + c;
            const ZCol zc = zc_of(min(col, NC - 1));
            const int2 zt = zst(zc);
            const double kk = MODE == Z_VORT ? kk_of(zc) : 0.0;
            if (col >= NC) return;
            // software pipelined like the Poisson epilogue: the w reads of the next PP row pairs
            // go out before the stores of the current ones
            struct WRaw { double c[PP][2], o[PP][2]; };
            auto w_load = [&](int ib) {
                WRaw r;
#pragma unroll
                for (int q = 0; q < PP; q++) {
                    const int i = min(ib + q * NPR, LE - 1);
#pragma unroll
                    for (int sd = 0; sd < 2; sd++) {
                        const size_t o = size_t(sd == 0 ? i : nz - 1 - i) * H2;
                        r.c[q][sd] = MODE == Z_VORT ? w_in[o + zc.so] : 0.0;
                        r.o[q][sd] = MODE == Z_VORT ? w_in[o + (zc.so ^ 1)] : 0.0;   // the other part of w
                    }
                }
                return r;
            };
            auto y_store = [&](int ib, const WRaw& r) {
#pragma unroll
                for (int q = 0; q < PP; q++) {
                    const int i = ib + q * NPR;
                    if (i >= LE) continue;
                    const double2 y = unfold_pair(G0, G1, nz, i, c);
#pragma unroll
                    for (int sd = 0; sd < 2; sd++) {
                        if (sd == 1 && i >= hF) continue;   // the centre row of odd nz is one node
                        const double yv = sd == 0 ? y.x : y.y;
                        const size_t ro = size_t(sd == 0 ? i : nz - 1 - i) * H2;
                        put(ro, zt, MODE == Z_PLAIN ? yv : vort(yv, r.c[q][sd], r.o[q][sd], kk, col & 1));
                    }
                }
            };
            constexpr int STEP = PP * NPR;
            const int i0 = threadIdx.x / BCOLS;
            WRaw ra = w_load(i0);
            for (int ib = i0; ib < LE; ib += 2 * STEP) {
                const WRaw rb = w_load(ib + STEP);   // (rows clamped into range)
                y_store(ib, ra);
                if (ib + STEP >= LE) break;
                ra = w_load(ib + 2 * STEP);
                y_store(ib + STEP, rb);
            }
            return;
        }
        if (a.d1.extra) {
            const int col = cb * BCOLS + int(threadIdx.x);
            if (threadIdx.x < BCOLS && col < NC) {
                const ZCol zc = zc_of(col);
                const size_t ro = size_t(hF) * H2;
                const double v = MODE == Z_PLAIN ? xs
                                 : vort(xs, w_in[ro + zc.so], w_in[ro + (zc.so ^ 1)], kk_of(zc), col & 1);
                put(ro, zst(zc), v);
            }
        }
        // per tile and row-block pair: the vorticity's w reads for all of the lane's nodes go out
        // together (indices clamped into range), then the results are combined and stored
#pragma unroll
        for (int ct = 0; ct < CT; ct++) {
            const int col = out_col(ct, cb), part = col & 1;
            const ZCol zc = zc_of(min(col, NC - 1));
            const int2 zt = zst(zc);
            const double kk = MODE == Z_VORT ? kk_of(zc) : 0.0;
#pragma unroll
            for (int ii = 0; ii < RBW; ii++) {
                const int b = wave + NW * ii;
                double yv[4][2], wc[4][2], wo[4][2];
                size_t qn[4][2];
                bool ok[4][2];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = out_row<NW, T>(ii, r);
                    const double sv = acc0[ii][ct][r];
                    const double rv = (row < hF && b < a.d1.rb1) ? double(acc1[ii][ct][r]) : 0.0;
                    yv[r][0] = sv + rv;
                    yv[r][1] = sv - rv;
#pragma unroll
                    for (int side = 0; side < 2; side++) {
                        const int node = max(min(side == 0 ? row : nz - 1 - row, nz - 1), 0);
                        qn[r][side] = size_t(node) * H2;
                        ok[r][side] = col < NC && b < a.d1.rb0 && row < LE && (side == 0 || row < hF);
                        if (MODE == Z_VORT) {
                            wc[r][side] = w_in[qn[r][side] + zc.so];
                            wo[r][side] = w_in[qn[r][side] + (zc.so ^ 1)];   // the other part of w
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int side = 0; side < 2; side++) {
                        if (!ok[r][side]) continue;
                        const double y = yv[r][side];
                        put(qn[r][side], zt, MODE == Z_PLAIN ? y : vort(y, wc[r][side], wo[r][side], kk, part));
                    }
            }
        }
    }
}

// F32: the FP32 contraction (the vorticity of the FP32-contraction precision)
template <int RBW, int MODE, int NW, bool F32 = false>
__global__ __launch_bounds__(64 * NW, Z_WAVES(RBW, ZW_D1)) void k_zderiv(ZArgs a) {
    extern __shared__ double smem[];
    using T = std::conditional_t<F32, float, double>;
    if (NW == 4 && a.packed && (zflags(a.zm, blockIdx.x) & ZB_PAIRED)) zderiv_tile<RBW, MODE, NW, NW == 4, T>(a, blockIdx.x, blockIdx.y, blockIdx.z, smem);
    else zderiv_tile<RBW, MODE, NW, false, T>(a, blockIdx.x, blockIdx.y, blockIdx.z, smem);
}
template <int RBW, int MODE, int NW, bool F32 = false>
__global__ __launch_bounds__(64 * NW, Z_WAVES(RBW, ZW_D1)) void k_zderiv_rag(const ZArgs* __restrict__ d, const Tile* __restrict__ map) {
    extern __shared__ double smem[];
    using T = std::conditional_t<F32, float, double>;
    const Tile t = map[blockIdx.x];
    const ZArgs& a = d[t.g];
    if (NW == 4 && a.packed && (zflags(a.zm, t.x) & ZB_PAIRED)) zderiv_tile<RBW, MODE, NW, NW == 4, T>(a, t.x, blockIdx.y, t.z, smem);
    else zderiv_tile<RBW, MODE, NW, false, T>(a, t.x, blockIdx.y, t.z, smem);
}

// apply_flow_control_forcing / update_pressure_gradient_3d / calculate_bulk_velocity_3d
// (DNS_pressure_BC_3D.f90:1107-1148, 2434-2549), called with istart as in :1118.
// U_b = sum_k w_k Re u_hat(0,0,k) / (nx ny sum w): the (0,0) coefficient of the
// unnormalised forward transform is the plane sum.  One wavefront (threads 0-63 of the block).
__device__ __forceinline__ void flow_tile(const FlowArgs& a, int mb) {
    const int lane = threadIdx.x;
    const FlowMember P = gmem(a.mem)[mb];
    const double* uh = gmem(a.uh) + size_t(mb) * a.mstride;
    const double* wz = gmem(a.wz);
    // the bulk integral is also formed under constant forcing: a non-finite value there flags
    // a member whose fields have gone non-finite (any such point reaches mode (0,0) of u
    // through the product's forward transform within one step)
    double part = 0.0;
    for (int k = lane; k < a.nz; k += 64) part += uh[size_t(k) * a.H2] * wz[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    if (lane != 0) return;
    FlowState s = gmem(a.st)[mb];
    if (P.method == 1) {
        s.G = P.target_G;
        s.Ub = part / (a.nxny * a.wsum);
    } else if (P.istart % P.freq == 0) {
        s.Ub = part / (a.nxny * a.wsum);
        s.err = P.target_Ub - s.Ub;
        const double kp = P.gain, ki = 0.05 * P.gain;
        const double gu = 3.0 / P.re + kp * s.err + ki * s.err_int;
        const double gc = fmax(0.0, fmin(10.0, gu));
        if (fabs(gc - gu) < 1.0e-12) {
            if (P.istart <= P.freq) s.err_int = s.err_int + s.err * a.dt;
            else s.err_int = s.err_int + 0.5 * a.dt * (s.err_prev + s.err);
        }
        s.err_prev = s.err;
        s.G = gc;
        s.gsum += s.G;
        s.gcount += 1;
        s.gavg = s.gsum / double(s.gcount);
    }
    gmem(a.st)[mb] = s;
}

// mode (0,0): row/column 1 of the Poisson matrix replaced by the identity, rhs_1 = 0
// (DNS_pressure_BC_3D.f90:1702-1711).  Its solution is linear in w*(0,0,:):
// phi = E0 w*, D1 phi = F0 w*, with E0 = [0; M0 W_{1:} D1_{1:,:}] / dt and F0 = D1 E0
// precomputed on the host.  Row k by one wavefront (threads 0-63 of the block).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void mode0_tile(const Mode0Args& a, int k, int mb) {
    const int nz = a.nz, lane = threadIdx.x;
    const size_t mo = size_t(mb) * a.mstride;
    const double* ws = gmem(a.ws) + mo;
    const double* e0 = gmem(a.e0);
    const double* f0 = gmem(a.f0);
    double s0 = 0.0, s1 = 0.0, d0 = 0.0, d1 = 0.0;
    for (int q = lane; q < nz; q += 64) {
        const double w0 = ws[size_t(q) * a.H2], w1 = ws[size_t(q) * a.H2 + 1];
        const double e = e0[size_t(k) * nz + q], f = f0[size_t(k) * nz + q];
        s0 += e * w0; s1 += e * w1; d0 += f * w0; d1 += f * w1;
    }
    s0 = wave_sum(s0); s1 = wave_sum(s1); d0 = wave_sum(d0); d1 = wave_sum(d1);
    if (lane < 2) {
        const int part = lane;
        const double xv = part ? s1 : s0, dx = part ? d1 : d0;
        const size_t o = mo + size_t(k) * a.H2 + part;
        if (!a.fused) {
            gmem(a.phi)[o] = xv;
            gmem(a.dphi)[o] = dx;
        } else {
            const bool wall = k == 0 || k == nz - 1;
            gmem(a.u)[o] = wall ? 0.0 : gmem(a.us)[o];
            gmem(a.v)[o] = wall ? 0.0 : gmem(a.vs)[o];
            gmem(a.w)[o] = wall ? 0.0 : ws[size_t(k) * a.H2 + part] - a.dt * dx;
            gmem(a.p)[o] = gmem(a.p)[o] + xv;
        }
    }
}

// -------------------------------------------------------------------------------------
// Helmholtz, one velocity component per grid.y (solve_viscous_helmholtz,
// DNS_pressure_BC_3D.f90:1425-1521, with the explicit half of Crank-Nicolson from
// add_explicit_diffusion :1355-1381 and the forcing of :1107-1148 folded in):
//   interior y = T (alpha u_I + dt S'_I) + f (T D2_I) u + c1 bc1 + c2 bc2,
//   z = y / (hl + f k^2), u*_I = B z, u*_0 = bc1, u*_{nz-1} = bc2,
// alpha = 1 - f k^2 (CN) or 1, S' = S + G nx ny at mode (0,0) of the u component, and
// bc = -dt (i k_eff p_hat) at the walls (the Kim & Moin values -dt grad p, :1325-1328).
// Folded: even eigen rows on [v_E ; u_E], odd rows on [v_O ; u_O] (v = alpha u_I + dt S'_I over
// the interior, u over the whole column); u*_I unfolds from p = B_E z_E and q = B_O z_O.
// -------------------------------------------------------------------------------------
// PS: the block holds a paired slot (its two parts are the real parts of two modes)
template <int RBW, bool SF32, int NW, bool PS>
__device__ __forceinline__ void helmholtz_tile(const HelmArgs& a, int cb, int comp, int mb, double* smem) {
    constexpr bool LX = NW < 4;
    constexpr int NT = 64 * NW;
    const int H2 = a.H2, nz = a.nz, m = a.m, P = a.zm.P, NC = 2 * P;
    const int hI = m / 2, L1 = hI + (m & 1), hF = nz / 2, LE = hF + (nz & 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, cl = lane & 15;
    const size_t mo = size_t(mb) * a.mstride, rs = size_t(H2) / 2;
    const double* uh = gmem(a.uh[comp]) + mo;
    const double* ph = gmem(a.ph) + mo;
    const double* xw = gmem(a.xw);
    const double* yw = gmem(a.yw);
    FoldOp F = a.hF;
    F.A += size_t(mb) * a.fstride;
    const double* c1v = gmem(a.c1) + size_t(mb) * m;
    const double* c2v = gmem(a.c2) + size_t(mb) * m;
    const double* hlv = gmem(a.hl) + size_t(mb) * m;
    const double fco = gmem(a.fvec)[mb];
    // the forcing enters mode (0,0) of u only, in the column block holding that mode,
    // component 0: that block runs the flow controller first when it is fused
    const bool forced = comp == 0 && cb == 0;   // slot 0 holds mode (0,0)
    const int bfl = NW == 4 ? zflags(a.zm, cb) : ZB_IDENT;   // (four-wave variants only)
    const bool sec = (bfl & ZB_SECOND) != 0, ident = (bfl & ZB_IDENT) != 0;
    if (a.flow_fused && forced) {
        if (threadIdx.x < 64) flow_tile(a.fl, mb);
        __syncthreads();
    }
    const double force = forced ? gmem(a.flow)[mb].G * a.nxny : 0.0;
    // staging thread: its slot's constants, once (alpha per part: the parts of a paired slot
    // are different modes)
    const int mt = threadIdx.x & 15, slot = cb * BMODES + mt;
    const bool sval = slot < P;
    const ZCol zx = zcol_t(a.zm, 2 * min(slot, P - 1), ident), zy = zcol_t(a.zm, 2 * min(slot, P - 1) + 1, ident);
    double ax = 1.0, ay = 1.0, fr = 0.0;
    if (sval) {
        const double kxx = xw[zx.i], kyx = yw[zx.j], kxy = xw[zy.i], kyy = yw[zy.j];
        ax = a.cn ? 1.0 - fco * (kxx * kxx + kyx * kyx) : 1.0;
        ay = a.cn ? 1.0 - fco * (kxy * kxy + kyy * kyy) : 1.0;
        fr = zx.so == 0 ? force : 0.0;
    }
    const int smode = sval ? (zx.so >> 1) : 0;   // storage mode of a plain slot
    const int ox = sval ? zx.so : 0, oy = sval ? zy.so : 0;   // storage columns of the two parts
    const double2* u2 = reinterpret_cast<const double2*>(uh) + smode;
    const float2* sf2 = reinterpret_cast<const float2*>(reinterpret_cast<const float*>(gmem(a.sh[comp])) + mo) + smode;
    const double2* sd2 = reinterpret_cast<const double2*>(gmem(a.sh[comp]) + mo) + smode;
    // Panel rows, branch-free: v rows fold v = alpha u_I + dt S'_I over the interior (v row c:
    // nodes c+1 and m-c), u rows fold u over the column (u row c: nodes c and nz-1-c); on a
    // centre row the two nodes coincide.  v row c and u row c+1 fold the same two nodes, so the
    // panel interleaves them, [v_0, u_1, v_1, u_2, ..., v_{L1-1}, u_L1, u_0] (LE = L1 + 1; the
    // host stores the operator's columns in this order): the two rows of a pair are staged by
    // lanes of one wave, whose load instruction then reads u once for both.  fetch reads u and S
    // at both nodes; on u rows, where S is not used, the S reads go to the u elements just read
    // (the same cache lines, no HBM traffic).
    using SV = typename std::conditional<SF32, float2, double2>::type;
    const SV* sv2 = SF32 ? reinterpret_cast<const SV*>(sf2) : reinterpret_cast<const SV*>(sd2);
    struct Raw { double2 ua, ub; SV sa, sb; };
    auto prow = [&](int kf, bool& inv, int& c) {   // panel row kf: a v row (inv) or a u row, index c
        const bool pz = kf < 2 * L1;
        inv = pz && !(kf & 1);
        c = pz ? (kf + 1) >> 1 : (kf == 2 * L1 ? 0 : LE);   // (rows past K: c = LE, kept zero)
    };
    auto nodes = [&](int kf, int& na, int& nb) {
        bool inv;
        int c;
        prow(kf, inv, c);
        na = min(inv ? c + 1 : c, nz - 1);
        nb = max(inv ? m - c : nz - 1 - c, 0);
    };
    auto fetch = [&](int kf) -> Raw {
        int na, nb;
        nodes(kf, na, nb);
        const bool inv = kf < 2 * L1 && !(kf & 1);
        if constexpr (PS) {
            // two scalar reads per value: a paired slot's parts are the real parts of two modes
            using ST = typename std::conditional<SF32, float, double>::type;
            const ST* s1 = reinterpret_cast<const ST*>(gmem(a.sh[comp])) + mo;
            auto ldu = [&](int n) { return make_double2(uh[size_t(n) * H2 + ox], uh[size_t(n) * H2 + oy]); };
            auto lsv = [&](int n) -> SV {
                if (!inv) return SV{};
                SV v;
                v.x = s1[size_t(n) * H2 + ox];
                v.y = s1[size_t(n) * H2 + oy];
                return v;
            };
            return Raw{ldu(na), ldu(nb), lsv(na), lsv(nb)};
        } else {
            const SV* sa = inv ? &sv2[na * rs] : reinterpret_cast<const SV*>(&u2[na * rs]);
            const SV* sb = inv ? &sv2[nb * rs] : reinterpret_cast<const SV*>(&u2[nb * rs]);
            return Raw{u2[na * rs], u2[nb * rs], *sa, *sb};
        }
    };
    auto finish = [&](int kf, const Raw& rw) -> FoldPair {
        bool inv;
        int c;
        prow(kf, inv, c);
        const int half = inv ? hI : hF, lim = inv ? L1 : LE;
        const double2 ua = rw.ua, ub = rw.ub;
        const double2 sa = make_double2(double(rw.sa.x), double(rw.sa.y));
        const double2 sb = make_double2(double(rw.sb.x), double(rw.sb.y));
        const double va_r = inv ? ax * ua.x + a.dt * (sa.x + fr) : ua.x;
        const double va_i = inv ? ay * ua.y + a.dt * sa.y : ua.y;
        const double vb_r = inv ? ax * ub.x + a.dt * (sb.x + fr) : ub.x;
        const double vb_i = inv ? ay * ub.y + a.dt * sb.y : ub.y;
        const double keep = (sval && c < lim) ? 1.0 : 0.0;
        const double pair = c < half ? 1.0 : 0.0;
        FoldPair r;
        r.p0 = make_double2(keep * (va_r + pair * vb_r), keep * (va_i + pair * vb_i));
        r.p1 = make_double2(keep * pair * (va_r - vb_r), keep * pair * (va_i - vb_i));
        return r;
    };
    const double xb = xrow_fetch(a.hB);   // the back operator's extra-row coefficient (parked below)
    d4 acc0[RBW][CT], acc1[RBW][CT];
    zero_acc2(acc0);
    zero_acc2(acc1);
    XAcc<NW> xacc;
    fold_gemm<RBW, false, NW>(acc0, acc1, F, fetch, finish, smem, xacc);
    const double xv = F.extra ? reduce_extra<NW>(xacc, smem) : 0.0;
    // per-lane column constants for the two 16-column tiles (reads clamped into range and
    // issued together; a column past the plane gets zeros), formed after the forward contraction
    // so that its prefetch registers are free of them
    double k2c[CT], bc1[CT], bc2[CT];
    const bool zwg = gmem(a.flow)[mb].zero_wallgrad != 0;   // grad p of the walls still zero (restart)
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = out_col(ct, cb), colc = min(col, NC - 1);
        const ZCol zc = zcol_t(a.zm, colc, ident);
        const int mode = zc.so >> 1, part = zc.so & 1, j = zc.j, i = zc.i;
        const double kx = xw[i], ky = yw[j];
        const double2 p0 = reinterpret_cast<const double2*>(ph)[mode];
        const double2 p1 = reinterpret_cast<const double2*>(ph + size_t(nz - 1) * H2)[mode];
        const double ke = comp == 0 ? kx_eff(xw, i, a.nx) : ky_eff(yw, i, j, a.nx, a.ny);
        const bool in = col < NC, lift = in && comp < 2 && !zwg;
        k2c[ct] = in ? kx * kx + ky * ky : 0.0;
        bc1[ct] = lift ? (part == 0 ? -a.dt * (-ke * p0.y) : -a.dt * (ke * p0.x)) : 0.0;   // -dt * (i ke p)
        bc2[ct] = lift ? (part == 0 ? -a.dt * (-ke * p1.y) : -a.dt * (ke * p1.x)) : 0.0;
    }
    const int ZR = zpanel_rows(a.hB.ks);
    double* z0 = smem;
    double* z1 = smem + ZR * LDB;
    xrow_park(a.hB, xb, z0);
    // the back contraction reads whole 16-row chunks: zero the panel rows past each group's rows
    // (every other row of a live column is written below; the barrier before the back
    // contraction orders these writes)
    zero_panel_tail<NW>(z0, F.rows0, ZR);
    zero_panel_tail<NW>(z1, F.rows1, ZR);
    if (F.extra && threadIdx.x < BCOLS && cb * BCOLS + int(threadIdx.x) < NC) {
        const int row = F.rows0 - 1, ct = threadIdx.x >> 4;
        z0[row * LDB + threadIdx.x] = (xv + c1v[row] * bc1[ct] + c2v[row] * bc2[ct]) / (hlv[row] + fco * k2c[ct]);
    }
    {
        // the lifting sums and eigenvalues of the lane's rows, read together (clamped) first
        double e1[2][RBW][4], e2[2][RBW][4], el[2][RBW][4];
#pragma unroll
        for (int ii = 0; ii < RBW; ii++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row<NW>(ii, r);
                const int g0 = min(row, F.rows0 - 1), g1 = F.rows0 + min(row, F.rows1 - 1);
                e1[0][ii][r] = c1v[g0]; e2[0][ii][r] = c2v[g0]; el[0][ii][r] = hlv[g0];
                e1[1][ii][r] = c1v[g1]; e2[1][ii][r] = c2v[g1]; el[1][ii][r] = hlv[g1];
            }
#pragma unroll
        for (int ct = 0; ct < CT; ct++) {
            const bool cv = out_col(ct, cb) < NC;
#pragma unroll
            for (int ii = 0; ii < RBW; ii++) {
                const int b = wave + NW * ii;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = out_row<NW>(ii, r);
                    if (cv && b < F.rb0 && row < F.rows0)
                        z0[row * LDB + ct * 16 + cl] = (acc0[ii][ct][r] + e1[0][ii][r] * bc1[ct] + e2[0][ii][r] * bc2[ct]) /
                                                       (el[0][ii][r] + fco * k2c[ct]);
                    if (cv && b < F.rb1 && row < F.rows1)
                        z1[row * LDB + ct * 16 + cl] = (acc1[ii][ct][r] + e1[1][ii][r] * bc1[ct] + e2[1][ii][r] * bc2[ct]) /
                                                       (el[1][ii][r] + fco * k2c[ct]);
                }
            }
        }
    }
    __syncthreads();
    zero_acc2(acc0);
    zero_acc2(acc1);
    fold_gemm_lds<RBW, NW>(acc0, acc1, a.hB, z0, z1);
    double* out = gmem(a.out[comp]) + mo;
    if constexpr (LX) {
        // few row blocks: unfold through LDS (over the z panels, once every wave has read them)
        const double xc = (a.hB.extra && threadIdx.x < BCOLS) ? extra_dot(a.hB, z0) : 0.0;
        __syncthreads();
        double* G0 = smem;
        double* G1 = smem + a.hB.rows0 * LDB;
        double* bcs = G1 + a.hB.rows1 * LDB;   // 2 x BCOLS wall values
        store_groups<RBW, NW>(acc0, acc1, a.hB, G0, G1);
        if (a.hB.extra && threadIdx.x < BCOLS) G0[(a.hB.rows0 - 1) * LDB + threadIdx.x] = xc;
        if (threadIdx.x < 16) {
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                bcs[ct * 16 + threadIdx.x] = bc1[ct];
                bcs[BCOLS + ct * 16 + threadIdx.x] = bc2[ct];
            }
        }
        __syncthreads();
        // each thread keeps one column (NT is a multiple of BCOLS) and walks the interior row
        // pairs i = t / BCOLS, + NT / BCOLS, ...: nodes 1 + i and m - i from one pair of LDS reads;
        // the first BCOLS threads also store the wall values
        const int c = threadIdx.x % BCOLS, col = cb * BCOLS + c;
        const int so = zcol_t(a.zm, min(col, NC - 1), ident).so;   // (storage column: ident here)
        if (col < NC) {
            if (threadIdx.x < BCOLS) {
                out[so] = bcs[c];
                out[size_t(nz - 1) * H2 + so] = bcs[BCOLS + c];
            }
            for (int i = threadIdx.x / BCOLS; i < L1; i += NT / BCOLS) {
                const double2 v = unfold_pair(G0, G1, m, i, c);
                out[size_t(1 + i) * H2 + so] = v.x;
                if (i < hI) out[size_t(m - i) * H2 + so] = v.y;
            }
        }
        return;
    }
    if (a.hB.extra && threadIdx.x < BCOLS && cb * BCOLS + int(threadIdx.x) < NC)   // the interior centre row
        zstore(out, size_t(a.hB.rows0) * H2, zst(zcol_t(a.zm, cb * BCOLS + threadIdx.x, ident)), extra_dot(a.hB, z0), sec);
    int2 zo[CT];   // storage of the lane's output columns (read here, after the contractions)
#pragma unroll
    for (int ct = 0; ct < CT; ct++) zo[ct] = zst(zcol_t(a.zm, min(out_col(ct, cb), NC - 1), ident));
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = out_col(ct, cb);
        if (col >= NC) continue;
#pragma unroll
        for (int ii = 0; ii < RBW; ii++) {
            const int b = wave + NW * ii;
            if (b >= a.hB.rb0) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row<NW>(ii, r);
                if (row >= L1) continue;
                const double p = acc0[ii][ct][r];
                const double q = (row < hI && b < a.hB.rb1) ? acc1[ii][ct][r] : 0.0;
                zstore(out, size_t(row + 1) * H2, zo[ct], p + q, sec);
                if (row < hI) zstore(out, size_t(m - row) * H2, zo[ct], p - q, sec);
            }
        }
        if (wave == 0 && (lane >> 4) == 0) {
            zstore(out, 0, zo[ct], bc1[ct], sec);
            zstore(out, size_t(nz - 1) * H2, zo[ct], bc2[ct], sec);
        }
    }
}

template <int RBW, bool SF32, int NW>
__global__ __launch_bounds__(64 * NW, Z_WAVES(RBW, ZW_HELM)) void k_helmholtz(HelmArgs a) {
    extern __shared__ double smem[];
    if (NW == 4 && (zflags(a.zm, blockIdx.x) & ZB_PAIRED)) helmholtz_tile<RBW, SF32, NW, NW == 4>(a, blockIdx.x, blockIdx.y, blockIdx.z, smem);
    else helmholtz_tile<RBW, SF32, NW, false>(a, blockIdx.x, blockIdx.y, blockIdx.z, smem);
}
template <int RBW, bool SF32, int NW>
__global__ __launch_bounds__(64 * NW, Z_WAVES(RBW, ZW_HELM)) void k_helmholtz_rag(const HelmArgs* __restrict__ d, const Tile* __restrict__ map) {
    extern __shared__ double smem[];
    const Tile t = map[blockIdx.x];
    const HelmArgs& a = d[t.g];
    if (NW == 4 && (zflags(a.zm, t.x) & ZB_PAIRED)) helmholtz_tile<RBW, SF32, NW, NW == 4>(a, t.x, blockIdx.y, t.z, smem);
    else helmholtz_tile<RBW, SF32, NW, false>(a, t.x, blockIdx.y, t.z, smem);
}

// -------------------------------------------------------------------------------------
// Pressure (solve_pressure_3d_standard :1654-1733 with compute_divergence_3d :1858-1879)
// and, when fused, the projection + wall values (pressure_correction_3d :1903-1929,
// apply_boundary_conditions_3d :1953-1972):
//   y = Tf div + (Tf D1) w*,  div = i kx_eff u* + i ky_eff v*  (the weak form,
//   solve_pressure_3d_weak_form :1792-1803: div = -(i xw u* + i yw v*), D1 -> -W^{-1} D1^T W),
//   z = -y / (dt (lf + k^2)), phi = Bf z, D1 phi = (D1 Bf) z,
//   u = u* - dt i kx_eff phi, v = v* - dt i ky_eff phi, w = w* - dt D1 phi, p += phi.
// Folded: even rows on [div_E ; w_O], odd rows on [div_O ; w_E]; phi unfolds from
// (Bf_E z_E, Bf_O z_O), D1 phi from ((D1 Bf)_O z_O, (D1 Bf)_E z_E).  Mode (0,0) is solved by
// mode0_tile (poisson_tile runs it after the tile).
// -------------------------------------------------------------------------------------
template <int RBW, int NW, bool PS>
__device__ __forceinline__ void poisson_body(const PoisArgs& a, int cb, int mb, double* smem) {
    constexpr bool LX = NW < 4;
    constexpr int NT = 64 * NW;
    const int H2 = a.H2, nz = a.nz, P = a.zm.P, NC = 2 * P, hF = nz / 2, LE = hF + (nz & 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, cl = lane & 15;
    const size_t mo = size_t(mb) * a.mstride, rs = size_t(H2) / 2;
    const double* us = gmem(a.us) + mo; const double* vs = gmem(a.vs) + mo; const double* ws = gmem(a.ws) + mo;
    double* uo = gmem(a.u) + mo; double* vo = gmem(a.v) + mo; double* wo = gmem(a.w) + mo; double* po = gmem(a.p) + mo;
    double* phio = gmem(a.phi) + mo; double* dphio = gmem(a.dphi) + mo;
    const double* xw = gmem(a.xw); const double* yw = gmem(a.yw); const double* pl = gmem(a.pl);
    // staging thread's slot; storage mode (0,0) (mode0_tile solves it) is left out part by part
    // (in a paired slot it is the first part only).  The weak form's horizontal term
    // -(i xw u* + i yw v*) is taken on the effective wavenumbers: they differ from the raw ones
    // only where the term is anti-Hermitian in a real field's transform (the kx = nx/2 column,
    // and ky = ny/2 on the kx = 0 and nx/2 columns), which the reference's c2r of phi drops.
    const int bfl = NW == 4 ? zflags(a.zm, cb) : ZB_IDENT;   // (four-wave variants only)
    const bool sec = (bfl & ZB_SECOND) != 0, ident = (bfl & ZB_IDENT) != 0;
    const int mt = threadIdx.x & 15, slot = cb * BMODES + mt;
    const bool sval = slot < P;
    const ZCol zx = zcol_t(a.zm, 2 * min(slot, P - 1), ident), zy = zcol_t(a.zm, 2 * min(slot, P - 1) + 1, ident);
    const int kpm = ((sval && zx.so != 0) ? 1 : 0) | ((sval && zy.so != 1) ? 2 : 0);   // parts kept
    double skx = 0.0, sky = 0.0;
    if (sval) {
        const double sg = a.weak ? -1.0 : 1.0;
        skx = sg * kx_eff(xw, zx.i, a.nx);
        sky = sg * ky_eff(yw, zx.i, zx.j, a.nx, a.ny);
    }
    const int smode = sval ? (zx.so >> 1) : 0;
    const int ox = sval ? zx.so : 0, oy = sval ? zy.so : 0;
    const double2* us2 = reinterpret_cast<const double2*>(us) + smode;
    const double2* vs2 = reinterpret_cast<const double2*>(vs) + smode;
    const double2* ws2 = reinterpret_cast<const double2*>(ws) + smode;
    // branch-free: rows [0, LE) fold div (E: sum on panel 0, O: difference on panel 1), rows
    // [LE, 2 LE) fold w* the other way round (w_O on panel 0, w_E on panel 1).  fetch reads
    // (u*, v*) on div rows and (w*, w*) on w rows at both nodes.
    struct Raw { double2 pa, pb, qa, qb; };
    auto fetch = [&](int kf) -> Raw {
        const bool ind = kf < LE;
        const int c = ind ? kf : kf - LE;
        const int na = min(c, nz - 1), nb = max(nz - 1 - c, 0);
        if constexpr (PS) {
            // a paired slot's parts are the real parts of two modes: two reads per value
            const double* p1 = ind ? us : ws;
            const double* q1 = ind ? vs : ws;
            auto ld = [&](const double* b, int n) { return make_double2(b[size_t(n) * H2 + ox], b[size_t(n) * H2 + oy]); };
            return Raw{ld(p1, na), ld(p1, nb), ld(q1, na), ld(q1, nb)};
        } else {
            const double2* p2 = ind ? us2 : ws2;
            const double2* q2 = ind ? vs2 : ws2;
            return Raw{p2[na * rs], p2[nb * rs], q2[na * rs], q2[nb * rs]};
        }
    };
    auto finish = [&](int kf, const Raw& rw) -> FoldPair {
        const bool ind = kf < LE;
        const int c = ind ? kf : kf - LE;
        const double2 ua = rw.pa, ub = rw.pb, va = rw.qa, vb = rw.qb;   // w rows: pa, pb = w*
        const double xa_r = ind ? -skx * ua.y - sky * va.y : ua.x, xa_i = ind ? skx * ua.x + sky * va.x : ua.y;
        const double xb_r = ind ? -skx * ub.y - sky * vb.y : ub.x, xb_i = ind ? skx * ub.x + sky * vb.x : ub.y;
        // the parts differ only in a paired slot
        const double kr = (c < LE && (kpm & 1)) ? 1.0 : 0.0, ki = PS ? ((c < LE && (kpm & 2)) ? 1.0 : 0.0) : kr;
        const double pair = c < hF ? 1.0 : 0.0;
        const double er = kr * (xa_r + pair * xb_r), ei = ki * (xa_i + pair * xb_i);   // even fold
        const double orr = kr * pair * (xa_r - xb_r), oi = ki * pair * (xa_i - xb_i);  // odd fold
        FoldPair r;
        r.p0 = ind ? make_double2(er, ei) : make_double2(orr, oi);
        r.p1 = ind ? make_double2(orr, oi) : make_double2(er, ei);
        return r;
    };
    // the back operators' extra-row coefficients (parked below: pB's beside z0, pDB's beside z1)
    const double xpb = xrow_fetch(a.pB), xpdb = xrow_fetch(a.pDB);
    d4 acc0[RBW][CT], acc1[RBW][CT];
    zero_acc2(acc0);
    zero_acc2(acc1);
    XAcc<NW> xacc;
    fold_gemm<RBW, false, NW>(acc0, acc1, a.pF, fetch, finish, smem, xacc);
    const double xv = a.pF.extra ? reduce_extra<NW>(xacc, smem) : 0.0;
    // per-lane column constants (after the forward contraction: its prefetch registers are free of them)
    double k2c[CT], kxe[CT], kye[CT];
    bool live[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = out_col(ct, cb);
        const ZCol zc = zcol_t(a.zm, min(col, NC - 1), ident);
        const int j = zc.j, i = zc.i;
        live[ct] = col < NC && zc.so >= 2;   // not storage mode (0,0)
        k2c[ct] = 1.0; kxe[ct] = 0.0; kye[ct] = 0.0;
        if (col < NC) {
            const double kx = xw[i], ky = yw[j];
            k2c[ct] = kx * kx + ky * ky;
            kxe[ct] = kx_eff(xw, i, a.nx);
            kye[ct] = ky_eff(yw, i, j, a.nx, a.ny);
        }
    }
    const int ZR = zpanel_rows(a.pB.ks);
    double* z0 = smem;
    double* z1 = smem + ZR * LDB;
    xrow_park(a.pB, xpb, z0);
    xrow_park(a.pDB, xpdb, z1);
    zero_panel_tail<NW>(z0, a.pF.rows0, ZR);
    zero_panel_tail<NW>(z1, a.pF.rows1, ZR);
    if (a.pF.extra && threadIdx.x < BCOLS) {
        const int row = a.pF.rows0 - 1, ct = threadIdx.x >> 4;
        if (live[ct]) z0[row * LDB + threadIdx.x] = -xv / (a.dt * (pl[row] + k2c[ct]));
    }
    {
        // eigenvalues of the lane's rows, read together (clamped) before the stores
        double l0[RBW][4], l1[RBW][4];
#pragma unroll
        for (int ii = 0; ii < RBW; ii++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row<NW>(ii, r);
                l0[ii][r] = pl[min(row, a.pF.rows0 - 1)];
                l1[ii][r] = pl[a.pF.rows0 + min(row, a.pF.rows1 - 1)];
            }
#pragma unroll
        for (int ct = 0; ct < CT; ct++) {
#pragma unroll
            for (int ii = 0; ii < RBW; ii++) {
                const int b = wave + NW * ii;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = out_row<NW>(ii, r);
                    if (live[ct] && b < a.pF.rb0 && row < a.pF.rows0)
                        z0[row * LDB + ct * 16 + cl] = -acc0[ii][ct][r] / (a.dt * (l0[ii][r] + k2c[ct]));
                    if (live[ct] && b < a.pF.rb1 && row < a.pF.rows1)
                        z1[row * LDB + ct * 16 + cl] = -acc1[ii][ct][r] / (a.dt * (l1[ii][r] + k2c[ct]));
                }
            }
        }
    }
    __syncthreads();
    // phi = Bf z: pressure increment and the horizontal projection
    zero_acc2(acc0);
    zero_acc2(acc1);
    fold_gemm_lds<RBW, NW>(acc0, acc1, a.pB, z0, z1);
    if constexpr (LX) {
        // few row blocks: both back contractions (phi and D1 phi) run while the z panels are
        // live, then both projections unfold through LDS over the dead panels (which lets more
        // blocks share a CU than unfold buffers of their own would)
        d4 dacc0[RBW][CT], dacc1[RBW][CT];
        zero_acc2(dacc0);
        zero_acc2(dacc1);
        fold_gemm_lds<RBW, NW>(dacc0, dacc1, a.pDB, z1, z0);
        const double xph = (a.pB.extra && threadIdx.x < BCOLS) ? extra_dot(a.pB, z0) : 0.0;
        const double xd = (a.pDB.extra && threadIdx.x < BCOLS) ? extra_dot(a.pDB, z1) : 0.0;
        __syncthreads();
        double* G0 = smem;
        double* G1 = G0 + a.pB.rows0 * LDB;
        double* cc = G1 + a.pB.rows1 * LDB;   // per block column: kx_eff, ky_eff, live
        if (threadIdx.x < 16) {
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                cc[ct * 16 + threadIdx.x] = kxe[ct];
                cc[BCOLS + ct * 16 + threadIdx.x] = kye[ct];
                cc[2 * BCOLS + ct * 16 + threadIdx.x] = live[ct] ? 1.0 : 0.0;
            }
        }
        store_groups<RBW, NW>(acc0, acc1, a.pB, G0, G1);
        if (a.pB.extra && threadIdx.x < BCOLS) G0[(a.pB.rows0 - 1) * LDB + threadIdx.x] = xph;
        __syncthreads();
        // Each thread keeps one column and walks row pairs i = t / BCOLS, + NPR, ...: nodes i and
        // nz-1-i unfold from one pair of LDS reads (branch free, unfold_pair), the other part of
        // the mode comes from the neighbouring lane (columns 2q, 2q+1 are lanes t, t^1), and the
        // reads of u*, v*, p for PP row pairs go out before any of their arithmetic and stores.
        constexpr int NPR = NT / BCOLS, PP = POIS_PP;
        const bool fused = a.fused != 0;
        const double dt = a.dt;
        const int c = threadIdx.x % BCOLS, col = cb * BCOLS + c, part = col & 1;
        const bool okc = col < NC && cc[2 * BCOLS + c] != 0.0;
        const double kxc = cc[c], kyc = cc[BCOLS + c];
        const int so = zcol_t(a.zm, min(col, NC - 1), ident).so;   // (storage column: ident here)
        const int i0 = threadIdx.x / BCOLS;
        // Software pipelined: the reads of the next PP row pairs are issued before the stores of
        // the current ones (vmcnt counts reads and stores in issue order, so a read issued after
        // a store would wait for the store too): ragged Poisson at nz = 65 246 -> 227 us.
        struct EpiRaw { double u[PP][2], v[PP][2], p[PP][2]; };
        auto epi_load = [&](int ib) {
            EpiRaw r;
#pragma unroll
            for (int q = 0; q < PP; q++) {
                const int i = min(ib + q * NPR, LE - 1);
#pragma unroll
                for (int sd = 0; sd < 2; sd++) {
                    // (read unconditionally, also when only phi is written: a conditional read
                    // would make the next set's reads wait for this set's)
                    const size_t o = size_t(sd == 0 ? i : nz - 1 - i) * H2 + so;
                    r.u[q][sd] = us[o];
                    r.v[q][sd] = vs[o];
                    r.p[q][sd] = po[o];
                }
            }
            return r;
        };
        auto epi_store = [&](int ib, const EpiRaw& r) {
#pragma unroll
            for (int q = 0; q < PP; q++) {
                const int i = ib + q * NPR;
                const double2 ph = unfold_pair(G0, G1, nz, min(i, LE - 1), c);
                const double2 ot = make_double2(__shfl_xor(ph.x, 1), __shfl_xor(ph.y, 1));
                if (i >= LE || !okc) continue;
#pragma unroll
                for (int sd = 0; sd < 2; sd++) {
                    if (sd == 1 && i >= hF) continue;   // the centre row of odd nz is one node
                    const int node = sd == 0 ? i : nz - 1 - i;
                    const double phv = sd == 0 ? ph.x : ph.y, other = sd == 0 ? ot.x : ot.y;
                    const size_t ro = size_t(node) * H2 + so;
                    if (!fused) {
                        phio[ro] = phv;
                        continue;
                    }
                    const double pre = part == 0 ? phv : other, pim = part == 0 ? other : phv;
                    const bool wall = node == 0 || node == nz - 1;
                    const double du = part == 0 ? -kxc * pim : kxc * pre;
                    const double dv = part == 0 ? -kyc * pim : kyc * pre;
                    uo[ro] = wall ? 0.0 : r.u[q][sd] - dt * du;
                    vo[ro] = wall ? 0.0 : r.v[q][sd] - dt * dv;
                    po[ro] = r.p[q][sd] + phv;
                }
            }
        };
        constexpr int STEP = PP * NPR;
        {
            EpiRaw ra = epi_load(i0);
            for (int ib = i0; ib < LE; ib += 2 * STEP) {
                const EpiRaw rb = epi_load(ib + STEP);   // (rows clamped into range)
                epi_store(ib, ra);
                if (ib + STEP >= LE) break;
                ra = epi_load(ib + 2 * STEP);
                epi_store(ib + STEP, rb);
            }
        }
        __syncthreads();
        store_groups<RBW, NW>(dacc0, dacc1, a.pDB, G0, G1);
        if (a.pDB.extra && threadIdx.x < BCOLS) G0[(a.pDB.rows0 - 1) * LDB + threadIdx.x] = xd;
        __syncthreads();
        struct WRaw { double w[PP][2]; };
        auto w_load = [&](int ib) {
            WRaw r;
#pragma unroll
            for (int q = 0; q < PP; q++) {
                const int i = min(ib + q * NPR, LE - 1);
#pragma unroll
                for (int sd = 0; sd < 2; sd++) r.w[q][sd] = ws[size_t(sd == 0 ? i : nz - 1 - i) * H2 + so];
            }
            return r;
        };
        auto w_store = [&](int ib, const WRaw& r) {
#pragma unroll
            for (int q = 0; q < PP; q++) {
                const int i = ib + q * NPR;
                if (i >= LE || !okc) continue;
                const double2 dph = unfold_pair(G0, G1, nz, i, c);
#pragma unroll
                for (int sd = 0; sd < 2; sd++) {
                    if (sd == 1 && i >= hF) continue;
                    const int node = sd == 0 ? i : nz - 1 - i;
                    const double d = sd == 0 ? dph.x : dph.y;
                    const size_t ro = size_t(node) * H2 + so;
                    if (!fused) dphio[ro] = d;
                    else wo[ro] = (node == 0 || node == nz - 1) ? 0.0 : r.w[q][sd] - dt * d;
                }
            }
        };
        {
            WRaw ra = w_load(i0);
            for (int ib = i0; ib < LE; ib += 2 * STEP) {
                const WRaw rb = w_load(ib + STEP);
                w_store(ib, ra);
                if (ib + STEP >= LE) break;
                ra = w_load(ib + 2 * STEP);
                w_store(ib + STEP, rb);
            }
        }
        return;
    }
    // phi epilogue, per 16-column tile: the outputs' old values (u*, v*, p) are read for all of
    // the lane's nodes first (indices clamped into range), then combined and stored
    int2 zo[CT];   // storage of the lane's output columns (read here, after the contractions)
#pragma unroll
    for (int ct = 0; ct < CT; ct++) zo[ct] = zst(zcol_t(a.zm, min(out_col(ct, cb), NC - 1), ident));
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = out_col(ct, cb);
        const int part = col & 1;
#pragma unroll
        for (int ii = 0; ii < RBW; ii++) {
            const int b = wave + NW * ii;
            double pv[4][2], ov[4][2];
            size_t qn[4][2];
            bool ok[4][2];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row<NW>(ii, r);
                const double p = acc0[ii][ct][r];
                const double q = (row < hF && b < a.pB.rb1) ? acc1[ii][ct][r] : 0.0;
                pv[r][0] = p + q;
                pv[r][1] = p - q;
                ov[r][0] = __shfl_xor(pv[r][0], 1);   // the other part of the mode
                ov[r][1] = __shfl_xor(pv[r][1], 1);
#pragma unroll
                for (int side = 0; side < 2; side++) {
                    const int node = min(side == 0 ? row : nz - 1 - row, nz - 1);
                    qn[r][side] = size_t(max(node, 0)) * H2 + zo[ct].x;
                    ok[r][side] = b < a.pB.rb0 && live[ct] && row < LE && (side == 0 || row < hF);
                }
            }
            if (!a.fused) {
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int side = 0; side < 2; side++)
                        if (ok[r][side]) zstore_q(phio, qn[r][side], zo[ct], pv[r][side], sec);
                continue;
            }
            double uu[4][2], vv[4][2], pp[4][2];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int side = 0; side < 2; side++) {
                    uu[r][side] = us[qn[r][side]];
                    vv[r][side] = vs[qn[r][side]];
                    pp[r][side] = po[qn[r][side]];
                }
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int side = 0; side < 2; side++) {
                    if (!ok[r][side]) continue;
                    const int row = out_row<NW>(ii, r);
                    const int node = side == 0 ? row : nz - 1 - row;
                    const double ph = pv[r][side], other = ov[r][side];
                    const double pre = part == 0 ? ph : other, pim = part == 0 ? other : ph;
                    const bool wall = node == 0 || node == nz - 1;
                    // u - dt*(i k phi): Re = u.re - dt*(-k phi.im), Im = u.im - dt*(k phi.re)
                    const double du = part == 0 ? -kxe[ct] * pim : kxe[ct] * pre;
                    const double dv = part == 0 ? -kye[ct] * pim : kye[ct] * pre;
                    const size_t q = qn[r][side];
                    zstore_q(uo, q, zo[ct], wall ? 0.0 : uu[r][side] - a.dt * du, sec);
                    zstore_q(vo, q, zo[ct], wall ? 0.0 : vv[r][side] - a.dt * dv, sec);
                    zstore_q(po, q, zo[ct], pp[r][side] + ph, sec);
                }
        }
    }
    if (a.pB.extra && wave == 0) {
        // the centre row phi_h (odd part zero there); lanes t and t^1 hold the two parts of a mode
        const int t = threadIdx.x, ct = (t >> 4) & 1;
        const double ph = t < BCOLS ? extra_dot(a.pB, z0) : 0.0;
        const double other = __shfl_xor(ph, 1);
        const int col = cb * BCOLS + t, node = a.pB.rows0 - 1;
        if (t < BCOLS && live[ct]) {   // lane t's column of tile ct is col: zo[ct] is its storage
            const size_t qn = size_t(node) * H2 + zo[ct].x;
            if (!a.fused) {
                zstore_q(phio, qn, zo[ct], ph, sec);
            } else {
                const int part = col & 1;
                const double pre = part == 0 ? ph : other, pim = part == 0 ? other : ph;
                const double du = part == 0 ? -kxe[ct] * pim : kxe[ct] * pre;
                const double dv = part == 0 ? -kye[ct] * pim : kye[ct] * pre;
                zstore_q(uo, qn, zo[ct], us[qn] - a.dt * du, sec);
                zstore_q(vo, qn, zo[ct], vs[qn] - a.dt * dv, sec);
                zstore_q(po, qn, zo[ct], po[qn] + ph, sec);
            }
        }
    }
    // D1 phi = (D1 Bf) z: the wall-normal projection (s on z_O, r on z_E)
    zero_acc2(acc0);
    zero_acc2(acc1);
    fold_gemm_lds<RBW, NW>(acc0, acc1, a.pDB, z1, z0);
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
#pragma unroll
        for (int ii = 0; ii < RBW; ii++) {
            const int b = wave + NW * ii;
            double dv[4][2], wv[4][2];
            size_t qn[4][2];
            bool ok[4][2];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row<NW>(ii, r);
                const double sv = acc0[ii][ct][r];
                const double rv = (row < hF && b < a.pDB.rb1) ? acc1[ii][ct][r] : 0.0;
                dv[r][0] = sv + rv;
                dv[r][1] = sv - rv;
#pragma unroll
                for (int side = 0; side < 2; side++) {
                    const int node = min(side == 0 ? row : nz - 1 - row, nz - 1);
                    qn[r][side] = size_t(max(node, 0)) * H2 + zo[ct].x;
                    ok[r][side] = b < a.pDB.rb0 && live[ct] && row < LE && (side == 0 || row < hF);
                    wv[r][side] = a.fused ? ws[qn[r][side]] : 0.0;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int side = 0; side < 2; side++) {
                    if (!ok[r][side]) continue;
                    const int row = out_row<NW>(ii, r);
                    const int node = side == 0 ? row : nz - 1 - row;
                    const size_t q = qn[r][side];
                    if (!a.fused) {
                        zstore_q(dphio, q, zo[ct], dv[r][side], sec);
                        continue;
                    }
                    const bool wall = node == 0 || node == nz - 1;
                    zstore_q(wo, q, zo[ct], wall ? 0.0 : wv[r][side] - a.dt * dv[r][side], sec);
                }
        }
    }
    if (a.pDB.extra && threadIdx.x < BCOLS) {
        // the centre row of D1 phi (s_h, on z_O)
        const int t = threadIdx.x, ct = t >> 4, node = a.pDB.rows0 - 1;
        if (live[ct]) {   // lane t's column of tile ct: zo[ct] is its storage
            const double dph = extra_dot(a.pDB, z1);
            const size_t qn = size_t(node) * H2 + zo[ct].x;
            if (!a.fused) zstore_q(dphio, qn, zo[ct], dph, sec);
            else zstore_q(wo, qn, zo[ct], ws[qn] - a.dt * dph, sec);
        }
    }
}

// the Poisson tile, then (when fused) wave 0 solves mode (0,0) at nodes cb, cb + ncb, ...
template <int RBW, int NW>
__device__ __forceinline__ void poisson_tile(const PoisArgs& a, int cb, int mb, double* smem) {
    if (NW == 4 && (zflags(a.zm, cb) & ZB_PAIRED)) poisson_body<RBW, NW, NW == 4>(a, cb, mb, smem);
    else poisson_body<RBW, NW, false>(a, cb, mb, smem);
    if (a.mode0 && threadIdx.x < 64) {
        const int ncb = zmap_blocks(a.zm);
        for (int k = cb; k < a.nz; k += ncb) mode0_tile(a.m0, k, mb);
    }
}

template <int RBW, int NW>
__global__ __launch_bounds__(64 * NW, Z_WAVES(RBW, ZW_POIS)) void k_poisson(PoisArgs a) {
    extern __shared__ double smem[];
    poisson_tile<RBW, NW>(a, blockIdx.x, blockIdx.z, smem);
}
// The same four-wave kernel compiled for two waves per SIMD (up to 256 VGPRs: it spills 32
// bytes instead of 128).  Launched when the grid fits two blocks per CU, where the third slot of
// the three-wave build would idle anyway: at 128 x 128 (512 column blocks) Poisson 71.4 -> 69.8 us
// FP64, 67.0 -> 63.8 us mixed.
template <int RBW, int NW>
__global__ __launch_bounds__(64 * NW, 2) void k_poisson_w2(PoisArgs a) {
    extern __shared__ double smem[];
    poisson_tile<RBW, NW>(a, blockIdx.x, blockIdx.z, smem);
}
static int g_z_cus = 0;   // compute units of the device (z_prepare)
template <int RBW, int NW>
__global__ __launch_bounds__(64 * NW, Z_WAVES(RBW, ZW_POIS)) void k_poisson_rag(const PoisArgs* __restrict__ d, const Tile* __restrict__ map) {
    extern __shared__ double smem[];
    const Tile t = map[blockIdx.x];
    const PoisArgs& a = d[t.g];
    poisson_tile<RBW, NW>(a, t.x, t.z, smem);
}

// projection of the stage API (pressure_correction_3d :1903-1929), from stored phi, D1 phi
__global__ __launch_bounds__(256) void k_correction(CorrArgs a) {
    const size_t total = size_t(a.nz) * (a.H2 / 2);
    const int nxh = a.nxh;
    for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < total; e += size_t(gridDim.x) * blockDim.x) {
        const size_t mode_all = e;
        const int mode = int(mode_all % (a.H2 / 2));
        const int j = mode / nxh, i = mode - j * nxh;
        const double kxe = kx_eff(a.xw, i, a.nx), kye = ky_eff(a.yw, i, j, a.nx, a.ny);
        const double2 ph = reinterpret_cast<const double2*>(a.phi)[e];
        const double2 dp = reinterpret_cast<const double2*>(a.dphi)[e];
        const double2 us = reinterpret_cast<const double2*>(a.us)[e];
        const double2 vs = reinterpret_cast<const double2*>(a.vs)[e];
        const double2 ws = reinterpret_cast<const double2*>(a.ws)[e];
        reinterpret_cast<double2*>(a.u)[e] = make_double2(us.x - a.dt * (-kxe * ph.y), us.y - a.dt * (kxe * ph.x));
        reinterpret_cast<double2*>(a.v)[e] = make_double2(vs.x - a.dt * (-kye * ph.y), vs.y - a.dt * (kye * ph.x));
        reinterpret_cast<double2*>(a.w)[e] = make_double2(ws.x - a.dt * dp.x, ws.y - a.dt * dp.y);
        double2 p = reinterpret_cast<double2*>(a.p)[e];
        reinterpret_cast<double2*>(a.p)[e] = make_double2(p.x + ph.x, p.y + ph.y);
    }
}

__global__ __launch_bounds__(256) void k_zero_walls(int nz, int H2, long long mstride, double* u, double* v, double* w) {
    const size_t mo = size_t(blockIdx.y) * mstride;
    u += mo; v += mo; w += mo;
    const size_t top = size_t(nz - 1) * H2;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < H2; c += gridDim.x * blockDim.x) {
        u[c] = 0.0; v[c] = 0.0; w[c] = 0.0;
        u[top + c] = 0.0; v[top + c] = 0.0; w[top + c] = 0.0;
    }
}

__global__ __launch_bounds__(64) void k_flow_control(FlowArgs a) { flow_tile(a, blockIdx.x); }
__global__ __launch_bounds__(256) void k_rk_stage(RKStageArgs a) {
    for (long long e = blockIdx.x * 256LL + threadIdx.x; e < a.n; e += 256LL * gridDim.x) {
        const long long plane = (e / a.H2) % a.nz;
        const bool wall = plane == 0 || plane == a.nz - 1;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double sv = a.s_f32 ? double(static_cast<const float*>(a.s[c])[e]) : static_cast<const double*>(a.s[c])[e];
            a.out[c][e] = wall ? 0.0 : a.un[c][e] + a.a * sv;
        }
    }
}

__device__ __forceinline__ double read_lane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// one wavefront per real column; lane owns rows lane + 64 r (nz <= 192)
__global__ __launch_bounds__(256) void k_jacgs(JacgsArgs a) {
    constexpr int R = 3;
    constexpr double small = 1.0e-30;
    const int lane = threadIdx.x & 63, c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= a.H2) return;
    const int nz = a.nz, mode = c >> 1, part = c & 1;
    const int j = mode / a.nxh, i = mode - j * a.nxh;
    const bool m0 = mode == 0;
    const size_t mo = size_t(blockIdx.y) * a.mstride;
    const double kxe = kx_eff(a.xw, i, a.nx), kye = ky_eff(a.yw, i, j, a.nx, a.ny);
    const double ksq = a.xw[i] * a.xw[i] + a.yw[j] * a.yw[j];
    double x[R], r[R], p[R], q[R], dg[R], ad[R], ap[R];
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
        const int row = lane + 64 * rr;
        x[rr] = r[rr] = p[rr] = q[rr] = dg[rr] = ad[rr] = 0.0;
        if (row >= nz) continue;
        const double wk = a.wz[row];
        // diagonal of A, as assembled (:1688-1691; mode (0,0): unit entry in row 1)
        ad[rr] = (m0 && row == 0) ? 1.0 : -a.st[size_t(row) * nz + row] - ksq * wk * (a.ybar / 2.0);
        dg[rr] = fabs(ad[rr]) > small ? 1.0 / ad[rr] : 1.0;
        const size_t o = mo + size_t(row) * a.H2 + 2 * size_t(mode);
        const double2 u = *reinterpret_cast<const double2*>(a.us + o);
        const double2 v = *reinterpret_cast<const double2*>(a.vs + o);
        const double2 dw = *reinterpret_cast<const double2*>(a.dws + o);
        const double div = part == 0 ? (-kxe * u.y) + (-kye * v.y) + dw.x : (kxe * u.x) + (kye * v.x) + dw.y;
        const double b = (m0 && row == 0) ? 0.0 : (div / a.dt) * (0.5 * a.ybar * wk);
        r[rr] = b;                       // b - A x0 with x0 = 0
        q[rr] = r[rr] * dg[rr];
        p[rr] = q[rr];
    }
    auto rsum = [&](const double (&u)[R], const double (&v)[R]) {
        double t = 0.0;
#pragma unroll
        for (int rr = 0; rr < R; rr++) t += u[rr] * v[rr];
        return wave_sum(t);
    };
    // A p: rows owned by the lane, columns in order (apn(i) = sum_j a(i,j) p(j))
    auto matvec = [&]() {
#pragma unroll
        for (int rr = 0; rr < R; rr++) ap[rr] = 0.0;
#pragma unroll
        for (int jj = 0; jj < R; jj++) {
            const int lim = min(64, nz - 64 * jj);
            for (int l = 0; l < lim; l++) {
                const int col = 64 * jj + l;
                const double pj = read_lane(p[jj], l);
                const double* sc = a.st + size_t(col) * nz;   // column col (= row col: symmetric)
#pragma unroll
                for (int rr = 0; rr < R; rr++) {
                    const int row = lane + 64 * rr;
                    if (row >= nz) continue;
                    double aij = row == col ? ad[rr] : -sc[row];
                    if (m0 && (row == 0 || col == 0) && row != col) aij = 0.0;
                    ap[rr] = __dadd_rn(ap[rr], __dmul_rn(aij, pj));
                }
            }
        }
    };
    const double res0 = sqrt(rsum(r, r));
    int it = 0;
    bool conv = res0 <= a.tol;
    if (!conv) {
        for (it = 1; it <= a.niter; it++) {
            matvec();
            const double qdr = rsum(q, r), pdapn = rsum(p, ap);
            const double alfa = qdr / (pdapn + small);
#pragma unroll
            for (int rr = 0; rr < R; rr++) {
                x[rr] = x[rr] + alfa * p[rr];
                r[rr] = r[rr] - alfa * ap[rr];
                q[rr] = r[rr] * dg[rr];
            }
            const double qdrnew = rsum(q, r);
            const double beta = qdrnew / (qdr + small);
#pragma unroll
            for (int rr = 0; rr < R; rr++) p[rr] = r[rr] * dg[rr] + beta * p[rr];
            const double res1 = sqrt(rsum(r, r));
            if (res1 <= a.tol) { conv = true; break; }
        }
        if (it > a.niter) it = a.niter;
    }
    if (lane == 0) {
        atomicAdd(a.iters, (unsigned long long)it);
        if (!conv) atomicAdd(a.nonconv, 1);
    }
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
        const int row = lane + 64 * rr;
        if (row < nz) a.phi[mo + size_t(row) * a.H2 + c] = x[rr];
    }
}

__global__ void k_jacgs_clear(unsigned long long* iters, int* nonconv) {
    if (threadIdx.x == 0) { *iters = 0ULL; *nonconv = 0; }
}

__global__ __launch_bounds__(64) void k_clear_wallgrad(FlowState* st, int n) {
    const int b = blockIdx.x * 64 + threadIdx.x;
    if (b < n) st[b].zero_wallgrad = 0;
}

// explicit roundings (no contraction into fma) reproduce the host generator bit for bit
__global__ __launch_bounds__(256) void k_init_fields(InitArgs a) {
    const size_t N = size_t(a.npl) * a.nz;
    for (size_t n = blockIdx.x * size_t(blockDim.x) + threadIdx.x; n < N; n += size_t(gridDim.x) * blockDim.x) {
        const int k = int(n / size_t(a.npl));
        const size_t q = n - size_t(k) * a.npl;
        double u = 0.0, v = 0.0;
        if (k != 0 && k != a.nz - 1) {
            double up = 0.0, vp = 0.0;
            if (a.perturbed) {
                up = __dmul_rn(a.u2[q], a.wd[k]);
                vp = __dmul_rn(a.v2[q], a.wd[k]);
                if (a.scaled) { up = __dmul_rn(up, a.sf); vp = __dmul_rn(vp, a.sf); }
            }
            u = __dadd_rn(a.U[k], up);
            v = __dadd_rn(0.0, vp);
        }
        a.u[n] = u; a.v[n] = v; a.w[n] = 0.0;
    }
}

// partial reductions: [max|u|, max|v|, max|w|, sum 0.5 |u|^2 w_k, max|d|]
__global__ __launch_bounds__(256) void k_reduce_partial(ReduceArgs a) {
    __shared__ double red[5][256];
    double mu = 0, mv = 0, mw = 0, e = 0, md = 0;
    const size_t N = size_t(a.npl) * a.nz;
    for (size_t n = blockIdx.x * size_t(blockDim.x) + threadIdx.x; n < N; n += size_t(gridDim.x) * blockDim.x) {
        const int k = int(n / size_t(a.npl));
        double uu = 0, vv = 0, ww = 0;
        if (a.u) { uu = a.u[n]; mu = fmax(mu, fabs(uu)); }
        if (a.v) { vv = a.v[n]; mv = fmax(mv, fabs(vv)); }
        if (a.w) { ww = a.w[n]; mw = fmax(mw, fabs(ww)); }
        e += 0.5 * (uu * uu + vv * vv + ww * ww) * a.wz[k];
        if (a.d) md = fmax(md, fabs(a.d[n]));
    }
    const int t = threadIdx.x;
    red[0][t] = mu; red[1][t] = mv; red[2][t] = mw; red[3][t] = e; red[4][t] = md;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            red[0][t] = fmax(red[0][t], red[0][t + s]);
            red[1][t] = fmax(red[1][t], red[1][t + s]);
            red[2][t] = fmax(red[2][t], red[2][t + s]);
            red[3][t] = red[3][t] + red[3][t + s];
            red[4][t] = fmax(red[4][t], red[4][t + s]);
        }
        __syncthreads();
    }
    if (t < 5) a.partial[size_t(blockIdx.x) * 5 + t] = red[t][0];
}

__global__ void k_reduce_final(ReduceArgs a) {
    if (threadIdx.x != 0) return;
    double r[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < a.nblocks; b++) {
        const double* p = a.partial + size_t(b) * 5;
        r[0] = fmax(r[0], p[0]); r[1] = fmax(r[1], p[1]); r[2] = fmax(r[2], p[2]);
        r[3] += p[3]; r[4] = fmax(r[4], p[4]);
    }
    for (int i = 0; i < 5; i++) a.result[i] = r[i];
}

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------
static inline dim3 col_grid(int H2, int y) { return dim3((H2 + BCOLS - 1) / BCOLS, y); }

// Dynamic LDS of the contraction kernels: fold_gemm's staging (stage_lds of the variant's waves,
// which also holds reduce_extra's 512 doubles), the two z panels of the back contractions,
// and the LDS unfold of the few-wave variants (G0, G1 over the dead panels, plus per-column values).
size_t zderiv_lds_bytes(const ZArgs& a) {
    const int var = zd_variant(a);
    const size_t g = z_variant_lx(var) ? size_t(a.d1.rows0 + a.d1.rows1) * LDB : 0;
    return sizeof(double) * std::max<size_t>(stage_lds(z_variant_nw(var)), g);
}
size_t helm_lds_bytes(const HelmArgs& a) {
    const int var = helm_variant(a);
    const size_t z = size_t(2) * zpanel_rows(a.hB.ks) * LDB;
    const size_t g = z_variant_lx(var) ? size_t(a.hB.rows0 + a.hB.rows1) * LDB + 2 * BCOLS : 0;
    return sizeof(double) * std::max<size_t>(std::max<size_t>(stage_lds(z_variant_nw(var)), z), g);
}
size_t pois_lds_bytes(const PoisArgs& a) {
    const int var = pois_variant(a);
    const size_t z = size_t(2) * zpanel_rows(a.pB.ks) * LDB;
    const size_t g = z_variant_lx(var) ? size_t(a.pB.rows0 + a.pB.rows1) * LDB + 3 * BCOLS : 0;
    return sizeof(double) * std::max<size_t>(std::max<size_t>(stage_lds(z_variant_nw(var)), z), g);
}

// variants (z_variant): (pairs per wave, waves per block) = (1, 1), (1, 2), (1, 3), (1, 4), (2, 4)
#define Z_BLOCK(var) dim3((var) < 3 ? 64 * ((var) + 1) : 256)
// (with f32, the vorticity only: the FP32 contraction; the plain D1 stays FP64)
#define ZD_CASES(K, G, ...)                                                                          \
    if (f32) {                                                                                       \
        if (mode != Z_VORT) return hipErrorInvalidValue;                                             \
        switch (var) {                                                                               \
            case 0: hipLaunchKernelGGL((K<1, Z_VORT, 1, true>), G, Z_BLOCK(var), __VA_ARGS__); break; \
            case 1: hipLaunchKernelGGL((K<1, Z_VORT, 2, true>), G, Z_BLOCK(var), __VA_ARGS__); break; \
            case 2: hipLaunchKernelGGL((K<1, Z_VORT, 3, true>), G, Z_BLOCK(var), __VA_ARGS__); break; \
            case 3: hipLaunchKernelGGL((K<1, Z_VORT, 4, true>), G, Z_BLOCK(var), __VA_ARGS__); break; \
            case 4: hipLaunchKernelGGL((K<2, Z_VORT, 4, true>), G, Z_BLOCK(var), __VA_ARGS__); break; \
            default: return hipErrorInvalidValue;                                                    \
        }                                                                                            \
    } else                                                                                           \
    switch (var * 2 + (mode == Z_VORT)) {                                                            \
        case 0: hipLaunchKernelGGL((K<1, Z_PLAIN, 1>), G, Z_BLOCK(var), __VA_ARGS__); break;          \
        case 1: hipLaunchKernelGGL((K<1, Z_VORT, 1>), G, Z_BLOCK(var), __VA_ARGS__); break;           \
        case 2: hipLaunchKernelGGL((K<1, Z_PLAIN, 2>), G, Z_BLOCK(var), __VA_ARGS__); break;          \
        case 3: hipLaunchKernelGGL((K<1, Z_VORT, 2>), G, Z_BLOCK(var), __VA_ARGS__); break;           \
        case 4: hipLaunchKernelGGL((K<1, Z_PLAIN, 3>), G, Z_BLOCK(var), __VA_ARGS__); break;          \
        case 5: hipLaunchKernelGGL((K<1, Z_VORT, 3>), G, Z_BLOCK(var), __VA_ARGS__); break;           \
        case 6: hipLaunchKernelGGL((K<1, Z_PLAIN, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;          \
        case 7: hipLaunchKernelGGL((K<1, Z_VORT, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;           \
        case 8: hipLaunchKernelGGL((K<2, Z_PLAIN, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;          \
        case 9: hipLaunchKernelGGL((K<2, Z_VORT, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;           \
        default: return hipErrorInvalidValue;                                                        \
    }

hipError_t launch_zderiv(int mode, const ZArgs& a, int nf, int nm, hipStream_t s) {
    dim3 g = col_grid(a.H2, mode == Z_VORT ? 2 : nf);
    if (a.packed) g.x = zmap_blocks(a.zm);
    g.z = nm;
    const int var = zd_variant(a);
    const size_t lds = zderiv_lds_bytes(a);
    const bool f32 = mode == Z_VORT && a.d1_f32 != 0;
    ZD_CASES(k_zderiv, g, lds, s, a)
    return hipGetLastError();
}

#define HM_CASES(K, G, ...)                                                                          \
    switch (var * 2 + (sf32 ? 1 : 0)) {                                                              \
        case 0: hipLaunchKernelGGL((K<1, false, 1>), G, Z_BLOCK(var), __VA_ARGS__); break;            \
        case 1: hipLaunchKernelGGL((K<1, true, 1>), G, Z_BLOCK(var), __VA_ARGS__); break;             \
        case 2: hipLaunchKernelGGL((K<1, false, 2>), G, Z_BLOCK(var), __VA_ARGS__); break;            \
        case 3: hipLaunchKernelGGL((K<1, true, 2>), G, Z_BLOCK(var), __VA_ARGS__); break;             \
        case 4: hipLaunchKernelGGL((K<1, false, 3>), G, Z_BLOCK(var), __VA_ARGS__); break;            \
        case 5: hipLaunchKernelGGL((K<1, true, 3>), G, Z_BLOCK(var), __VA_ARGS__); break;             \
        case 6: hipLaunchKernelGGL((K<1, false, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;            \
        case 7: hipLaunchKernelGGL((K<1, true, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;             \
        case 8: hipLaunchKernelGGL((K<2, false, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;            \
        case 9: hipLaunchKernelGGL((K<2, true, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;             \
        default: return hipErrorInvalidValue;                                                        \
    }

hipError_t launch_helmholtz(const HelmArgs& a, int nm, hipStream_t s) {
    dim3 g(zmap_blocks(a.zm), 3);
    g.z = nm;
    const size_t lds = helm_lds_bytes(a);
    const int var = helm_variant(a);
    const bool sf32 = a.s_f32 != 0;
    HM_CASES(k_helmholtz, g, lds, s, a)
    return hipGetLastError();
}

#define PS_CASES(K, G, ...)                                                                          \
    switch (var) {                                                                                   \
        case 0: hipLaunchKernelGGL((K<1, 1>), G, Z_BLOCK(var), __VA_ARGS__); break;                   \
        case 1: hipLaunchKernelGGL((K<1, 2>), G, Z_BLOCK(var), __VA_ARGS__); break;                   \
        case 2: hipLaunchKernelGGL((K<1, 3>), G, Z_BLOCK(var), __VA_ARGS__); break;                   \
        case 3: hipLaunchKernelGGL((K<1, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;                   \
        case 4: hipLaunchKernelGGL((K<2, 4>), G, Z_BLOCK(var), __VA_ARGS__); break;                   \
        default: return hipErrorInvalidValue;                                                        \
    }

hipError_t launch_poisson(const PoisArgs& a, int nm, hipStream_t s) {
    dim3 g(zmap_blocks(a.zm), 1);
    g.z = nm;
    const size_t lds = pois_lds_bytes(a);
    const int var = pois_variant(a);
    if (var == 3 && g_z_cus > 0 && size_t(g.x) * g.z <= size_t(2) * g_z_cus) {
        hipLaunchKernelGGL((k_poisson_w2<1, 4>), g, Z_BLOCK(var), lds, s, a);
        return hipGetLastError();
    }
    PS_CASES(k_poisson, g, lds, s, a)
    return hipGetLastError();
}
hipError_t z_prepare() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&g_z_cus, hipDeviceAttributeMultiprocessorCount, dev);
    return e;
}
hipError_t launch_correction(const CorrArgs& a, hipStream_t s) {
    const size_t total = size_t(a.nz) * (a.H2 / 2);
    int blocks = int(std::min<size_t>((total + 255) / 256, 4096));
    hipLaunchKernelGGL(k_correction, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_zero_walls(int nz, int H2, long long mstride, int nm, double* u, double* v, double* w,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_zero_walls, dim3((H2 + 255) / 256, nm), dim3(256), 0, s, nz, H2, mstride, u, v, w);
    return hipGetLastError();
}
hipError_t launch_flow_control(const FlowArgs& a, int nm, hipStream_t s) {
    hipLaunchKernelGGL(k_flow_control, dim3(nm), dim3(64), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_rk_stage(const RKStageArgs& a, hipStream_t s) {
    const int blocks = int(std::min<long long>((a.n + 255) / 256, 8192));
    hipLaunchKernelGGL(k_rk_stage, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_jacgs(const JacgsArgs& a, int nm, hipStream_t s) {
    if (a.nz > 192) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_jacgs, dim3((a.H2 + 3) / 4, nm), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_jacgs_clear(unsigned long long* iters, int* nonconv, hipStream_t s) {
    hipLaunchKernelGGL(k_jacgs_clear, dim3(1), dim3(64), 0, s, iters, nonconv);
    return hipGetLastError();
}
hipError_t launch_clear_wallgrad(FlowState* st, int nm, hipStream_t s) {
    if (nm <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_wallgrad, dim3((nm + 63) / 64), dim3(64), 0, s, st, nm);
    return hipGetLastError();
}
hipError_t launch_zderiv_rag(int var, int mode, bool f32, const ZArgs* d, const Tile* map, int nt, int nf, size_t lds,
                             hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    const dim3 g(nt, mode == Z_VORT ? 2 : nf);
    ZD_CASES(k_zderiv_rag, g, lds, s, d, map)
    return hipGetLastError();
}
hipError_t launch_helmholtz_rag(int var, bool sf32, const HelmArgs* d, const Tile* map, int nt, size_t lds,
                                hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    const dim3 g(nt, 3);
    HM_CASES(k_helmholtz_rag, g, lds, s, d, map)
    return hipGetLastError();
}
hipError_t launch_poisson_rag(int var, const PoisArgs* d, const Tile* map, int nt, size_t lds, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    PS_CASES(k_poisson_rag, dim3(nt), lds, s, d, map)
    return hipGetLastError();
}
// plain copy of n doubles (calibration of the HBM byte counters)
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ x, double2* __restrict__ y, size_t n2) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n2; i += size_t(gridDim.x) * blockDim.x) y[i] = x[i];
}
hipError_t launch_copy(const double* x, double* y, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, reinterpret_cast<const double2*>(x),
                       reinterpret_cast<double2*>(y), n / 2);
    return hipGetLastError();
}

hipError_t launch_init_fields(const InitArgs& a, hipStream_t s) {
    const size_t N = size_t(a.npl) * a.nz;
    const int blocks = int(std::min<size_t>((N + 255) / 256, 8192));
    hipLaunchKernelGGL(k_init_fields, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partial, dim3(a.nblocks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace dns

namespace dns {
// Lane-layout probe of v_mfma_f64_16x16x4f64: block b uses a one-hot A operand at lane b/64
// and a one-hot B operand at lane b%64; out[b*256 + lane*4 + q] = result item q of lane.
__global__ __launch_bounds__(64) void k_mfma_probe(double* out) {
    const int lane = threadIdx.x;
    const int la = blockIdx.x / 64, lb = blockIdx.x % 64;
    const double a = lane == la ? 1.0 : 0.0, b = lane == lb ? 1.0 : 0.0;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    double* o = out + size_t(blockIdx.x) * 256 + lane * 4;
    o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3];
}
hipError_t launch_mfma_probe(double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(4096), dim3(64), 0, s, out);
    return hipGetLastError();
}

// Same probe for v_mfma_f32_16x16x4f32 (result items converted to double).
typedef float f4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(64) void k_mfma_probe_f32(double* out) {
    const int lane = threadIdx.x;
    const int la = blockIdx.x / 64, lb = blockIdx.x % 64;
    const float a = lane == la ? 1.0f : 0.0f, b = lane == lb ? 1.0f : 0.0f;
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    double* o = out + size_t(blockIdx.x) * 256 + lane * 4;
    o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3];
}
hipError_t launch_mfma_probe_f32(double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma_probe_f32, dim3(4096), dim3(64), 0, s, out);
    return hipGetLastError();
}
}  // namespace dns
