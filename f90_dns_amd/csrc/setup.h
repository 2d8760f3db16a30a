// setup.h -- host-side grid constants, LGL operators and per-mode solve bases.
//
// Everything here runs once per context on the host, in FP64 (the eigen bases in
// 80-bit long double, rounded once), and is uploaded to HBM.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dns {

// LGL nodes and weights, Newton on P'_{n-1} from -cos(pi (i-1)/(n-1)), tolerance 1e-15,
// at most 100 iterations (lgl_module.f90:30-80).
void lgl_nodes_weights(int n, std::vector<double>& x, std::vector<double>& w);

// Lagrange differentiation matrix, column-major D[i + n*j] = D(i,j), with the
// reference's rule that an entry is zero when |prod_{k!=j}(x_j - x_k)| <= 1e-15
// (lgl_module.f90:211-263).  For n >= 65 every off-diagonal entry is zero.
void differentiation_matrix(int n, const std::vector<double>& x, std::vector<double>& d);

// Exact LGL differentiation matrix (not the reference's; opt-in): D(i,j) = P_N(x_i) /
// (P_N(x_j)(x_i - x_j)), D(0,0) = -N(N+1)/4, D(N,N) = N(N+1)/4, N = n-1, in long double.
void differentiation_matrix_exact(int n, const std::vector<double>& x, std::vector<double>& d);

// Wavenumbers of setup_spectral_3d (DNS_pressure_BC_3D.f90:769-796).
void wavenumbers(int nx, int ny, double xlen, double ylen, std::vector<double>& xw, std::vector<double>& yw);

// 2/3 dealias mask of apply_dealias_truncation (DNS_pressure_BC_3D.f90:1163-1193), 1 = kept.
void dealias_mask(int nx, int ny, std::vector<unsigned char>& m);

// Matrix in the lane order of v_mfma_f64_16x16x4f64's A operand: for row block rb and
// k-step s, entry [(rb*KS + s)*64 + lane] = M(rb*16 + lane%16, 4*s + lane/16), zero outside
// the rows x cols of M.  M is given row-major (rows x cols).
struct SwizzledA {
    int rows = 0, cols = 0, rb = 0, ks = 0;
    std::vector<double> data;
};
SwizzledA swizzle_a(const std::vector<double>& rowmajor, int rows, int cols);

// Even/odd folding.  The LGL nodes are symmetric (z_{n-1-k} = -z_k), so D1 is
// centro-antisymmetric, D2 and the stiffness matrix are centrosymmetric, and the eigenbases of
// the per-mode solves are even or odd vectors.  A column vector x of length n = 2h + r
// (r = n mod 2) folds into X_E(c) = x_c + x_{n-1-c} (c < h) with X_E(h) = x_h when r = 1, and
// X_O(c) = x_c - x_{n-1-c} (c < h).  Every wall-normal operator then splits into two blocks of
// half size; an output that comes back in physical rows unfolds from its two groups g0 (h + r
// rows, the one holding the centre) and g1 (h rows) as out_i = g0_i + g1_i,
// out_{n-1-i} = g0_i - g1_i (i < h), out_h = g0_h.
//
// FoldedA: the two row groups in the lane order of v_mfma_f64_16x16x4f64's A operand; group g
// acts on its own K-row panel.  Block cb (group-0 blocks first, then group 1 from block pb),
// k-step s: entry [(cb*ksp + s)*64 + lane] = M_g(16*(cb - first_g) + lane%16, 4s + lane/16);
// ks = ceil(K/4) k-steps are used; ksp (a multiple of 8, >= ks + 1) are stored, the tail zero.
// When rows0 = 16 k + 1 = rows1 + 1 (the centre row or the extra even eigenvector of an odd
// column), that last group-0 row is not in the blocks (rb0 = k) but stored dense after them
// (extra = 1, 4 ksp entries) and computed as a dot product, so the row-block pairs stay k per
// group.  Even folds (rows0 = rows1, no centre row) keep every row in the blocks.
// Both groups are padded with zero blocks to pb = 4 ceil(max(rb0, rb1) / 4) blocks, so each of
// a block's 4 waves loads its pairs unconditionally; group 1 starts at block pb.
struct FoldedA {
    int rows0 = 0, rows1 = 0, K = 0, rb0 = 0, rb1 = 0, ks = 0, ksp = 0, extra = 0, pb = 0;
    std::vector<double> data;
};
FoldedA swizzle_fold(const std::vector<double>& M0, int rows0, const std::vector<double>& M1, int rows1, int K);
// Column of the Helmholtz forward operator [T | f T D2_I] (v rows 0..L1-1, then u rows 0..L1) that
// panel row p holds in the interleaved order [v_0, u_1, v_1, u_2, ..., v_{L1-1}, u_L1, u_0].
inline int helm_panel_column(int p, int L1) { return p < 2 * L1 ? ((p & 1) ? L1 + (p + 1) / 2 : p / 2) : L1; }

// Operators of one grid and run (nz, dt, Re, theta):
//  Helmholtz (interior unknowns I = 1..nz-2, m = nz-2), with Q L Q^T = W_I^{-1/2} A_II W_I^{-1/2}:
//    y = T (alpha u_I + dt S_I) + f*(T D2_I) u + c1*bc1 + c2*bc2,  z = y / (hl + f k^2),
//    x_I = B z, x_0 = bc1, x_{nz-1} = bc2,
//    T = Q^T W_I^{1/2}, B = W_I^{-1/2} Q, c_{1,2} = -2 f Q^T W_I^{-1/2} A_{I,{0,nz-1}}, hl = f lambda + 1.
//    Folded: eigenvectors ordered even (group 0), then odd (group 1); the forward operator acts on
//    the panels [v_E ; u_E] (group 0) and [v_O ; u_O] (group 1) with v = alpha u_I + dt S_I folded
//    over the interior (L1 rows) and u over all nz rows (L1 + 1 rows); B z unfolds (p from z_E, q
//    from z_O).  The stored K order interleaves the rows that fold the same two nodes (v row c and
//    u row c+1): [v_0, u_1, v_1, u_2, ..., v_{L1-1}, u_L1, u_0] (helm_panel_column).
//  Poisson (all nz unknowns), with Qf Lf Qf^T = W^{-1/2} A W^{-1/2}:
//    y = Tf div + (Tf D1) w,  z = -y / (dt (lf + k^2)),  phi = Bf z,  D1 phi = (D1 Bf) z,
//    Tf = Qf^T W^{1/2}, Bf = W^{-1/2} Qf.  Folded: group 0 (even) on [div_E ; w_O], group 1 (odd)
//    on [div_O ; w_E]; phi unfolds from (p = Bf_E z_E, q = Bf_O z_O), D1 phi from
//    (s = (D1 Bf)_O z_O, r = (D1 Bf)_E z_E).  Mode (0,0): phi_0 = 0, phi_{1:} = M0 b_{1:} with
//    M0 = -(A_{1:,1:})^{-1} and b = w * (D1 w*) / dt, i.e. phi = E0 w*, D1 phi = F0 w* with
//    E0 = [0; M0 W_{1:} D1_{1:,:}] / dt and F0 = D1 E0.  The weak-form RHS replaces D1 on w*
//    by Gw = -W^{-1} D1^T W (in Tf D1 and E0) and div by -(i kx u* + i ky v*) on the raw
//    wavenumbers.
//  z derivative: D1 x unfolds from (s = D1 x_odd on X_O, r = D1 x_even on X_E).
struct Operators {
    int nz = 0, m = 0;
    double f = 0.0;              // theta dt / Re
    std::vector<double> z, wz, d1, d2, stiff;     // d1, d2, stiff column-major nz x nz
    FoldedA zD1;                 // s: (h+r) rows on X_O | r: h rows on X_E, K = h + r
    FoldedA hF;                  // even | odd eigen rows on [v ; u] panels, K = mE + (h + r)
    FoldedA hB;                  // p | q rows on z_E | z_O, K = mE
    std::vector<double> hc1, hc2, hl;   // length m, even eigen rows first
    FoldedA pF;                  // even | odd eigen rows on [div_E ; w_O] | [div_O ; w_E]
    FoldedA pB;                  // p | q on z_E | z_O
    FoldedA pDB;                 // s | r on z_O | z_E
    std::vector<double> pl;      // length nz, even first
    std::vector<double> m0;      // (nz-1) x (nz-1) row-major
    std::vector<double> e0, f0;  // nz x nz row-major: mode-(0,0) phi = E0 w*, D1 phi = F0 w*
    double helm_recon_err = 0, pois_recon_err = 0;   // max |reconstruction - operator|, relative
};
// weak: the pressure RHS of the weak form (w* enters through -W^{-1} D1^T W instead of D1)
bool build_operators(int nz, double dt, double re, double theta, bool crank_nicolson, bool exact_d1, bool weak,
                     Operators& op, std::string& err);

// exp(-2 pi i k / n), k = 0..n-1, as interleaved (re, im)
void twiddles(int n, std::vector<double>& tw);

// Radix factors of n (4s, then 2, then 3); false if n has a prime factor > 3.
bool radix_factors(int n, std::vector<int>& f);

struct InitSpec {
    int nx, ny, nz;
    double xlen, ylen, u_centerline;
    int enable_perturbations, kind, mode_x, mode_y;
    double amplitude;
    uint64_t seed;
    int kmax;
};
// initialize_fields_3d + apply_solenoidal_perturbations (DNS_pressure_BC_3D.f90:909-962,3219-3265;
// perturbation_analytical_simple.f90:24-155); kind 1 = seeded stream-function noise.
void initial_field(const InitSpec& s, std::vector<double>& u, std::vector<double>& v, std::vector<double>& w);

// The same field in separable form, for assembly on the device: at interior plane k,
//   u = U[k] + scale(u2[q] * wd[k]),  v = scale(v2[q] * wd[k]),  w = 0,
// with scale(x) = x * sf when scaled (else x), every operation rounded as in initial_field;
// walls are zero.  u2, v2 are planar (nx*ny), U and wd have nz entries.
struct PlanarInit {
    std::vector<double> u2, v2, U, wd;
    double sf = 1.0;
    int scaled = 0, perturbed = 0;
};
void initial_planar(const InitSpec& s, PlanarInit& out);

}  // namespace dns
