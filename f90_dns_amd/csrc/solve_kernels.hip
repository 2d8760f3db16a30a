// solve_kernels.hip -- wall-normal (z) work on spectral columns: the LGL D1 contraction,
// the Crank-Nicolson Helmholtz solves and the pressure Poisson solve + projection, each as
// FP64 MFMA contractions over 16-column tiles per wavefront, plus the small kernels of the
// step (mode (0,0) Poisson, correction, wall values, PI forcing, reductions).
//
// Spectral fields are complex(nx/2+1, ny, nz), plane-major; seen as real matrices they are
// nz rows x H2 = 2 (nx/2+1) ny columns, column c = 2*mode + part (part 0 = Re, 1 = Im).
#include <algorithm>

#include "contract.h"
#include "kernels.h"

namespace dns {

static_assert(BMODES == 16, "staging assigns 16 modes x 16 rows to 256 threads");

// Wavenumber seen by a derivative whose result is projected by a c2r/r2c round trip:
// the x Nyquist column gives an anti-Hermitian column (zero after the round trip), and so
// does the y Nyquist row in the kx = 0 and kx = nx/2 columns.
__device__ __forceinline__ double kx_eff(const double* xw, int i, int nx) { return i == nx / 2 ? 0.0 : xw[i]; }
__device__ __forceinline__ double ky_eff(const double* yw, int i, int j, int nx, int ny) {
    return (j == ny / 2 && (i == 0 || i == nx / 2)) ? 0.0 : yw[j];
}

// -------------------------------------------------------------------------------------
// z derivative: out = D1 in  (compute_derivatives_3d dz, DNS_pressure_BC_3D.f90:2348-2358),
// or the vorticity components that need it: omega_x = i ky w - D1 v, omega_y = D1 u - i kx w
// (DNS_pressure_BC_3D.f90:1077-1079).  One block = 16 modes, all nz rows.
// -------------------------------------------------------------------------------------
// Tile (cb, by, mb): column block cb of member mb; by = field (Z_PLAIN).
template <int RBW, int MODE>
__device__ __forceinline__ void zderiv_tile(const ZArgs& a, int cb, int by, int mb, double* lds) {
    const int H2 = a.H2, nz = a.nz, H = H2 / 2, RB = (nz + 15) / 16;
    const int mode0 = cb * BMODES;
    const int nfield = MODE == Z_VORT ? 2 : 1;
    const size_t mo = size_t(mb) * a.mstride;
    const double* w_in = a.in[2] + mo;
    for (int q = 0; q < nfield; q++) {
        const int f = MODE == Z_VORT ? (q == 0 ? 1 : 0) : by;   // vort: v first, then u
        const double* in = a.in[f] + mo;
        d4 acc[RBW][CT];
        zero_acc2(acc);
        gemm_staged<RBW>(acc, a.Asw, RB, a.KS, [&](int k, int ml) -> double2 {
            const int mode = mode0 + ml;
            if (k < nz && mode < H) return reinterpret_cast<const double2*>(in + size_t(k) * H2)[mode];
            return make_double2(0.0, 0.0);
        }, lds);
#pragma unroll
        for (int ct = 0; ct < CT; ct++) {
            const int col = out_col(ct, cb);
            if (col >= H2) continue;
            const int mode = col >> 1, part = col & 1, j = mode / a.nxh, i = mode - j * a.nxh;
#pragma unroll
            for (int ii = 0; ii < RBW; ii++) {
#pragma unroll 2
                for (int r = 0; r < 4; r++) {
                    const int row = out_row(ii, r);
                    if (row >= nz) continue;
                    const double y = acc[ii][ct][r];
                    if (MODE == Z_PLAIN) {
                        a.out[f][mo + size_t(row) * H2 + col] = y;
                    } else {
                        double res;
                        const double wo = w_in[size_t(row) * H2 + (col ^ 1)];   // the other part of w
                        const double wr = part == 0 ? w_in[size_t(row) * H2 + col] : wo;
                        const double wi = part == 0 ? wo : w_in[size_t(row) * H2 + col];
                        if (f == 1) {        // omega_x = i ky w - D1 v
                            const double ky = a.yw[j];
                            res = part == 0 ? (-ky * wi - y) : (ky * wr - y);
                        } else {             // omega_y = D1 u - i kx w
                            const double kx = a.xw[i];
                            res = part == 0 ? (y + kx * wi) : (y - kx * wr);
                        }
                        const int o = f == 1 ? 0 : 1;
                        if (a.out_f32) reinterpret_cast<float*>(a.out[o])[mo + size_t(row) * H2 + col] = float(res);
                        else a.out[o][mo + size_t(row) * H2 + col] = res;
                    }
                }
            }
        }
    }
}

template <int RBW, int MODE>
__global__ __launch_bounds__(256) void k_zderiv(ZArgs a) {
    __shared__ double lds[2 * KC * LDB];
    zderiv_tile<RBW, MODE>(a, blockIdx.x, blockIdx.y, blockIdx.z, lds);
}
template <int RBW, int MODE>
__global__ __launch_bounds__(256) void k_zderiv_rag(const ZArgs* __restrict__ d, const Tile* __restrict__ map) {
    __shared__ double lds[2 * KC * LDB];
    const Tile t = map[blockIdx.x];
    const ZArgs& a = d[t.g];
    zderiv_tile<RBW, MODE>(a, t.x, blockIdx.y, t.z, lds);
}

// -------------------------------------------------------------------------------------
// Helmholtz, one velocity component per grid.y (solve_viscous_helmholtz,
// DNS_pressure_BC_3D.f90:1425-1521, with the explicit half of Crank-Nicolson from
// add_explicit_diffusion :1355-1381 and the forcing of :1107-1148 folded in):
//   interior y = T (alpha u_I + dt S'_I) + f (T D2_I) u + c1 bc1 + c2 bc2,
//   z = y / (hl + f k^2), u*_I = B z, u*_0 = bc1, u*_{nz-1} = bc2,
// alpha = 1 - f k^2 (CN) or 1, S' = S + G nx ny at mode (0,0) of the u component, and
// bc = -dt (i k_eff p_hat) at the walls (the Kim & Moin values -dt grad p, :1325-1328).
// -------------------------------------------------------------------------------------
template <int RBW>
constexpr int solve_lds() { return (4 * RBW * 16 > 2 * KC ? 4 * RBW * 16 : 2 * KC) * LDB; }

template <int RBW>
__device__ __forceinline__ void helmholtz_tile(const HelmArgs& a, int cb, int comp, int mb, double* smem) {
    const int H2 = a.H2, nz = a.nz, m = a.m, H = H2 / 2, RBI = (m + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mode0 = cb * BMODES;
    const size_t mo = size_t(mb) * a.mstride;
    const double* uh = a.uh[comp] + mo;
    const double* sh = a.s_f32 ? reinterpret_cast<const double*>(reinterpret_cast<const float*>(a.sh[comp]) + mo)
                               : a.sh[comp] + mo;
    const double* ph = a.ph + mo;
    const double* Fsw = a.Fsw + size_t(mb) * a.fstride;
    const double* c1v = a.c1 + size_t(mb) * m;
    const double* c2v = a.c2 + size_t(mb) * m;
    const double* hlv = a.hl + size_t(mb) * m;
    const double fco = a.fvec[mb];
    const double force = comp == 0 ? a.flow[mb].G * a.nxny : 0.0;
    // per-lane column constants for the two 16-column tiles
    double k2c[CT], bc1[CT], bc2[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = out_col(ct, cb);
        k2c[ct] = 0.0; bc1[ct] = 0.0; bc2[ct] = 0.0;
        if (col < H2) {
            const int mode = col >> 1, part = col & 1, j = mode / a.nxh, i = mode - j * a.nxh;
            const double kx = a.xw[i], ky = a.yw[j];
            k2c[ct] = kx * kx + ky * ky;
            if (comp < 2) {
                const double ke = comp == 0 ? kx_eff(a.xw, i, a.nx) : ky_eff(a.yw, i, j, a.nx, a.ny);
                const double2 p0 = reinterpret_cast<const double2*>(ph)[mode];
                const double2 p1 = reinterpret_cast<const double2*>(ph + size_t(nz - 1) * H2)[mode];
                bc1[ct] = part == 0 ? -a.dt * (-ke * p0.y) : -a.dt * (ke * p0.x);    // -dt * (i ke p)
                bc2[ct] = part == 0 ? -a.dt * (-ke * p1.y) : -a.dt * (ke * p1.x);
            }
        }
    }
    d4 acc[RBW][CT];
    zero_acc2(acc);
    gemm_staged<RBW>(acc, Fsw, RBI, a.FKS, [&](int k, int ml) -> double2 {
        const int mode = mode0 + ml;
        if (mode >= H) return make_double2(0.0, 0.0);
        if (k < m) {
            const int j = mode / a.nxh, i = mode - j * a.nxh;
            const double kx = a.xw[i], ky = a.yw[j];
            const double alpha = a.cn ? 1.0 - fco * (kx * kx + ky * ky) : 1.0;
            const double2 u = reinterpret_cast<const double2*>(uh + size_t(k + 1) * H2)[mode];
            double2 sv;
            if (a.s_f32) {
                const float2 sf = reinterpret_cast<const float2*>(reinterpret_cast<const float*>(sh) + size_t(k + 1) * H2)[mode];
                sv = make_double2(double(sf.x), double(sf.y));
            } else {
                sv = reinterpret_cast<const double2*>(sh + size_t(k + 1) * H2)[mode];
            }
            const double fr = mode == 0 ? force : 0.0;
            return make_double2(alpha * u.x + a.dt * (sv.x + fr), alpha * u.y + a.dt * sv.y);
        }
        if (k < m + nz) return reinterpret_cast<const double2*>(uh + size_t(k - m) * H2)[mode];
        return make_double2(0.0, 0.0);
    }, smem);
    double* zs = smem;
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
#pragma unroll
        for (int ii = 0; ii < RBW; ii++) {
            if (wave + 4 * ii >= RBI) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row(ii, r);
                double z = 0.0;
                if (row < m && out_col(ct, cb) < H2)
                    z = (acc[ii][ct][r] + c1v[row] * bc1[ct] + c2v[row] * bc2[ct]) / (hlv[row] + fco * k2c[ct]);
                zs[row * LDB + ct * 16 + (lane & 15)] = z;
            }
        }
    }
    __syncthreads();
    zero_acc2(acc);
    gemm_lds<RBW>(acc, a.Bsw, RBI, a.BKS, zs);
    double* out = a.out[comp] + mo;
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = out_col(ct, cb);
        if (col >= H2) continue;
#pragma unroll
        for (int ii = 0; ii < RBW; ii++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row(ii, r);
                if (wave + 4 * ii < RBI && row < m) out[size_t(row + 1) * H2 + col] = acc[ii][ct][r];
            }
        }
        if (wave == 0 && (lane >> 4) == 0) {
            out[col] = bc1[ct];
            out[size_t(nz - 1) * H2 + col] = bc2[ct];
        }
    }
}

template <int RBW>
__global__ __launch_bounds__(256) void k_helmholtz(HelmArgs a) {
    __shared__ double smem[solve_lds<RBW>()];
    helmholtz_tile<RBW>(a, blockIdx.x, blockIdx.y, blockIdx.z, smem);
}
template <int RBW>
__global__ __launch_bounds__(256) void k_helmholtz_rag(const HelmArgs* __restrict__ d, const Tile* __restrict__ map) {
    __shared__ double smem[solve_lds<RBW>()];
    const Tile t = map[blockIdx.x];
    const HelmArgs& a = d[t.g];
    helmholtz_tile<RBW>(a, t.x, blockIdx.y, t.z, smem);
}

// -------------------------------------------------------------------------------------
// Pressure (solve_pressure_3d_standard :1654-1733 with compute_divergence_3d :1858-1879)
// and, when fused, the projection + wall values (pressure_correction_3d :1903-1929,
// apply_boundary_conditions_3d :1953-1972):
//   y = Tf div + (Tf D1) w*,  div = i kx_eff u* + i ky_eff v*,
//   z = -y / (dt (lf + k^2)), phi = Bf z, D1 phi = (D1 Bf) z,
//   u = u* - dt i kx_eff phi, v = v* - dt i ky_eff phi, w = w* - dt D1 phi, p += phi.
// Mode (0,0) is left to k_poisson_mode0.
// -------------------------------------------------------------------------------------
template <int RBW>
__device__ __forceinline__ void poisson_tile(const PoisArgs& a, int cb, int mb, double* smem) {
    const int H2 = a.H2, nz = a.nz, H = H2 / 2, RB = (nz + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mode0 = cb * BMODES;
    const size_t mo = size_t(mb) * a.mstride;
    const double* us = a.us + mo; const double* vs = a.vs + mo; const double* ws = a.ws + mo;
    double* uo = a.u + mo; double* vo = a.v + mo; double* wo = a.w + mo; double* po = a.p + mo;
    double* phio = a.phi + mo; double* dphio = a.dphi + mo;
    double k2c[CT], kxe[CT], kye[CT];
    bool live[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = out_col(ct, cb);
        const int mode = col >> 1, j = mode / a.nxh, i = mode - j * a.nxh;
        live[ct] = col < H2 && mode != 0;
        k2c[ct] = 1.0; kxe[ct] = 0.0; kye[ct] = 0.0;
        if (col < H2) {
            const double kx = a.xw[i], ky = a.yw[j];
            k2c[ct] = kx * kx + ky * ky;
            kxe[ct] = kx_eff(a.xw, i, a.nx);
            kye[ct] = ky_eff(a.yw, i, j, a.nx, a.ny);
        }
    }
    d4 acc[RBW][CT];
    zero_acc2(acc);
    gemm_staged<RBW>(acc, a.Fsw, RB, a.FKS, [&](int k, int ml) -> double2 {
        const int mode = mode0 + ml;
        if (mode >= H || mode == 0) return make_double2(0.0, 0.0);
        if (k < nz) {
            const int j = mode / a.nxh, i = mode - j * a.nxh;
            const double ke = kx_eff(a.xw, i, a.nx), kf = ky_eff(a.yw, i, j, a.nx, a.ny);
            const double2 u = reinterpret_cast<const double2*>(us + size_t(k) * H2)[mode];
            const double2 v = reinterpret_cast<const double2*>(vs + size_t(k) * H2)[mode];
            return make_double2(-ke * u.y - kf * v.y, ke * u.x + kf * v.x);
        }
        if (k < 2 * nz) return reinterpret_cast<const double2*>(ws + size_t(k - nz) * H2)[mode];
        return make_double2(0.0, 0.0);
    }, smem);
    double* zs = smem;
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
#pragma unroll
        for (int ii = 0; ii < RBW; ii++) {
            if (wave + 4 * ii >= RB) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row(ii, r);
                double z = 0.0;
                if (row < nz && live[ct]) z = -acc[ii][ct][r] / (a.dt * (a.pl[row] + k2c[ct]));
                zs[row * LDB + ct * 16 + (lane & 15)] = z;
            }
        }
    }
    __syncthreads();
    // phi = Bf z: pressure increment and the horizontal projection
    zero_acc2(acc);
    gemm_lds<RBW>(acc, a.Bsw, RB, a.BKS, zs);
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = out_col(ct, cb);
        const int part = col & 1;
#pragma unroll
        for (int ii = 0; ii < RBW; ii++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row(ii, r);
                const double ph = acc[ii][ct][r];
                const double other = __shfl_xor(ph, 1);       // the other part of this mode
                if (!live[ct] || wave + 4 * ii >= RB || row >= nz) continue;
                const size_t q = size_t(row) * H2 + col;
                if (!a.fused) {
                    phio[q] = ph;
                    continue;
                }
                const double pre = part == 0 ? ph : other, pim = part == 0 ? other : ph;
                const bool wall = row == 0 || row == nz - 1;
                // u - dt*(i k phi): Re = u.re - dt*(-k phi.im), Im = u.im - dt*(k phi.re)
                const double du = part == 0 ? -kxe[ct] * pim : kxe[ct] * pre;
                const double dv = part == 0 ? -kye[ct] * pim : kye[ct] * pre;
                uo[q] = wall ? 0.0 : us[q] - a.dt * du;
                vo[q] = wall ? 0.0 : vs[q] - a.dt * dv;
                po[q] = po[q] + ph;
            }
        }
    }
    // D1 phi = (D1 Bf) z: the wall-normal projection
    zero_acc2(acc);
    gemm_lds<RBW>(acc, a.DBsw, RB, a.BKS, zs);
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = out_col(ct, cb);
#pragma unroll
        for (int ii = 0; ii < RBW; ii++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = out_row(ii, r);
                if (!live[ct] || wave + 4 * ii >= RB || row >= nz) continue;
                const size_t q = size_t(row) * H2 + col;
                if (!a.fused) {
                    dphio[q] = acc[ii][ct][r];
                    continue;
                }
                const bool wall = row == 0 || row == nz - 1;
                wo[q] = wall ? 0.0 : ws[q] - a.dt * acc[ii][ct][r];
            }
        }
    }
}

template <int RBW>
__global__ __launch_bounds__(256) void k_poisson(PoisArgs a) {
    __shared__ double smem[solve_lds<RBW>()];
    poisson_tile<RBW>(a, blockIdx.x, blockIdx.z, smem);
}
template <int RBW>
__global__ __launch_bounds__(256) void k_poisson_rag(const PoisArgs* __restrict__ d, const Tile* __restrict__ map) {
    __shared__ double smem[solve_lds<RBW>()];
    const Tile t = map[blockIdx.x];
    const PoisArgs& a = d[t.g];
    poisson_tile<RBW>(a, t.x, t.z, smem);
}

// mode (0,0): row/column 1 of the Poisson matrix replaced by the identity, rhs_1 = 0
// (DNS_pressure_BC_3D.f90:1702-1711).  Its solution is linear in w*(0,0,:):
// phi = E0 w*, D1 phi = F0 w*, with E0 = [0; M0 W_{1:} D1_{1:,:}] / dt and F0 = D1 E0
// precomputed on the host.  One block per row k, one wavefront.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void mode0_tile(const Mode0Args& a, int k, int mb) {
    const int nz = a.nz, lane = threadIdx.x;
    const size_t mo = size_t(mb) * a.mstride;
    const double* ws = a.ws + mo;
    double s0 = 0.0, s1 = 0.0, d0 = 0.0, d1 = 0.0;
    for (int q = lane; q < nz; q += 64) {
        const double w0 = ws[size_t(q) * a.H2], w1 = ws[size_t(q) * a.H2 + 1];
        const double e = a.e0[size_t(k) * nz + q], f = a.f0[size_t(k) * nz + q];
        s0 += e * w0; s1 += e * w1; d0 += f * w0; d1 += f * w1;
    }
    s0 = wave_sum(s0); s1 = wave_sum(s1); d0 = wave_sum(d0); d1 = wave_sum(d1);
    if (lane < 2) {
        const int part = lane;
        const double xv = part ? s1 : s0, dx = part ? d1 : d0;
        const size_t o = mo + size_t(k) * a.H2 + part;
        if (!a.fused) {
            a.phi[o] = xv;
            a.dphi[o] = dx;
        } else {
            const bool wall = k == 0 || k == nz - 1;
            a.u[o] = wall ? 0.0 : a.us[o];
            a.v[o] = wall ? 0.0 : a.vs[o];
            a.w[o] = wall ? 0.0 : a.ws[o] - a.dt * dx;
            a.p[o] = a.p[o] + xv;
        }
    }
}

__global__ __launch_bounds__(64) void k_poisson_mode0(Mode0Args a) { mode0_tile(a, blockIdx.x, blockIdx.y); }
__global__ __launch_bounds__(64) void k_poisson_mode0_rag(const Mode0Args* __restrict__ d, const Tile* __restrict__ map) {
    const Tile t = map[blockIdx.x];
    const Mode0Args& a = d[t.g];
    mode0_tile(a, t.x, t.z);
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

// apply_flow_control_forcing / update_pressure_gradient_3d / calculate_bulk_velocity_3d
// (DNS_pressure_BC_3D.f90:1107-1148, 2434-2549), called with istart as in :1118.
// U_b = sum_k w_k Re u_hat(0,0,k) / (nx ny sum w): the (0,0) coefficient of the
// unnormalised forward transform is the plane sum.
__device__ __forceinline__ void flow_tile(const FlowArgs& a, int mb) {
    const int lane = threadIdx.x;
    const FlowMember P = a.mem[mb];
    const double* uh = a.uh + size_t(mb) * a.mstride;
    // the bulk integral is also formed under constant forcing: a non-finite value there flags
    // a member whose fields have gone non-finite (any such point reaches mode (0,0) of u
    // through the product's forward transform within one step)
    double part = 0.0;
    for (int k = lane; k < a.nz; k += 64) part += uh[size_t(k) * a.H2] * a.wz[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    if (lane != 0) return;
    FlowState s = a.st[mb];
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
    a.st[mb] = s;
}

__global__ __launch_bounds__(64) void k_flow_control(FlowArgs a) { flow_tile(a, blockIdx.x); }
__global__ __launch_bounds__(64) void k_flow_control_rag(const FlowArgs* __restrict__ d, const Tile* __restrict__ map) {
    const Tile t = map[blockIdx.x];
    const FlowArgs& a = d[t.g];
    flow_tile(a, t.z);
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
static inline int rbw_of(int rows) { return ((rows + 15) / 16 + 3) / 4; }

hipError_t launch_zderiv(int rb, int mode, const ZArgs& a, int nf, int nm, hipStream_t s) {
    const int rbw = (rb + 3) / 4;
    dim3 g = col_grid(a.H2, mode == Z_VORT ? 1 : nf);
    g.z = nm;
    switch (rbw * 2 + (mode == Z_VORT)) {
        case 2: hipLaunchKernelGGL((k_zderiv<1, Z_PLAIN>), g, dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_zderiv<1, Z_VORT>), g, dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_zderiv<2, Z_PLAIN>), g, dim3(256), 0, s, a); break;
        case 5: hipLaunchKernelGGL((k_zderiv<2, Z_VORT>), g, dim3(256), 0, s, a); break;
        case 6: hipLaunchKernelGGL((k_zderiv<3, Z_PLAIN>), g, dim3(256), 0, s, a); break;
        case 7: hipLaunchKernelGGL((k_zderiv<3, Z_VORT>), g, dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_helmholtz(int rbi, const HelmArgs& a, int nm, hipStream_t s) {
    dim3 g = col_grid(a.H2, 3);
    g.z = nm;
    switch ((rbi + 3) / 4) {
        case 1: hipLaunchKernelGGL((k_helmholtz<1>), g, dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_helmholtz<2>), g, dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_helmholtz<3>), g, dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_poisson(int rb, const PoisArgs& a, int nm, hipStream_t s) {
    dim3 g = col_grid(a.H2, 1);
    g.z = nm;
    switch ((rb + 3) / 4) {
        case 1: hipLaunchKernelGGL((k_poisson<1>), g, dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_poisson<2>), g, dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_poisson<3>), g, dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_poisson_mode0(const Mode0Args& a, int nm, hipStream_t s) {
    hipLaunchKernelGGL(k_poisson_mode0, dim3(a.nz, nm), dim3(64), 0, s, a);
    return hipGetLastError();
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
hipError_t launch_zderiv_rag(int rbw, int mode, const ZArgs* d, const Tile* map, int nt, int nf, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    const dim3 g(nt, mode == Z_VORT ? 1 : nf);
    switch (rbw * 2 + (mode == Z_VORT)) {
        case 2: hipLaunchKernelGGL((k_zderiv_rag<1, Z_PLAIN>), g, dim3(256), 0, s, d, map); break;
        case 3: hipLaunchKernelGGL((k_zderiv_rag<1, Z_VORT>), g, dim3(256), 0, s, d, map); break;
        case 4: hipLaunchKernelGGL((k_zderiv_rag<2, Z_PLAIN>), g, dim3(256), 0, s, d, map); break;
        case 5: hipLaunchKernelGGL((k_zderiv_rag<2, Z_VORT>), g, dim3(256), 0, s, d, map); break;
        case 6: hipLaunchKernelGGL((k_zderiv_rag<3, Z_PLAIN>), g, dim3(256), 0, s, d, map); break;
        case 7: hipLaunchKernelGGL((k_zderiv_rag<3, Z_VORT>), g, dim3(256), 0, s, d, map); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_helmholtz_rag(int rbw, const HelmArgs* d, const Tile* map, int nt, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    const dim3 g(nt, 3);
    switch (rbw) {
        case 1: hipLaunchKernelGGL((k_helmholtz_rag<1>), g, dim3(256), 0, s, d, map); break;
        case 2: hipLaunchKernelGGL((k_helmholtz_rag<2>), g, dim3(256), 0, s, d, map); break;
        case 3: hipLaunchKernelGGL((k_helmholtz_rag<3>), g, dim3(256), 0, s, d, map); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_poisson_rag(int rbw, const PoisArgs* d, const Tile* map, int nt, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    switch (rbw) {
        case 1: hipLaunchKernelGGL((k_poisson_rag<1>), dim3(nt), dim3(256), 0, s, d, map); break;
        case 2: hipLaunchKernelGGL((k_poisson_rag<2>), dim3(nt), dim3(256), 0, s, d, map); break;
        case 3: hipLaunchKernelGGL((k_poisson_rag<3>), dim3(nt), dim3(256), 0, s, d, map); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_poisson_mode0_rag(const Mode0Args* d, const Tile* map, int nt, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_poisson_mode0_rag, dim3(nt), dim3(64), 0, s, d, map);
    return hipGetLastError();
}
hipError_t launch_flow_control_rag(const FlowArgs* d, const Tile* map, int nt, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_flow_control_rag, dim3(nt), dim3(64), 0, s, d, map);
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
