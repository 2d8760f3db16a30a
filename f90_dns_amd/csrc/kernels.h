// kernels.h -- argument blocks and host launchers of the device kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace dns {

// Ragged launches: one launch serves tiles of several contexts ("groups") of different shape.
// Block b of a ragged launch works on tile map[b] of group map[b].g, whose argument block is
// descs[map[b].g]; x = plane (plane kernels), column block (z-kernels) or row (mode (0,0)),
// z = member within the group.
struct Tile {
    int g, x, z;
};

// ---- plane FFT kernels ------------------------------------------------------------------
enum C2ROp { C2R_PLAIN = 0, C2R_NL = 1, C2R_DIV = 2 };
enum R2COp { R2C_PLAIN = 0, R2C_CROSS = 1 };

// Plane-size classes of the plane kernels: threads per block and the largest plane (nx/2)*ny
// complex values they take (eight per thread; eighteen in class 4, which fits the LDS only in
// FP32).  Class 0 is one wavefront.  Classes 5 and 6 (nine and twelve waves) take the planes of
// 4608 and 6144 values (e.g. 96 x 96, 128 x 96), which left a 1024-thread block a third or a
// quarter idle in every radix-8 stage.  In order of capacity: 0, 1, 2, 5, 6, 3, 4
// (fft_class_promote).
constexpr int FFT_CLASSES = 7;
constexpr int CLS_NT[FFT_CLASSES] = {64, 256, 512, 1024, 1024, 576, 768};
constexpr int CLS_MAXE[FFT_CLASSES] = {512, 2048, 4096, 8192, 18432, 4608, 6144};
// complex values a thread may hold in one transform stage (registers)
constexpr int stage_budget(bool f32) { return f32 ? 36 : 18; }
// butterflies per thread a radix-R stage of class cls is compiled for
constexpr int stage_maxb(int cls, int R, bool f32) {
    return ((CLS_MAXE[cls] / R + CLS_NT[cls] - 1) / CLS_NT[cls]) < (stage_budget(f32) / R > 1 ? stage_budget(f32) / R : 1)
               ? (CLS_MAXE[cls] / R + CLS_NT[cls] - 1) / CLS_NT[cls]
               : (stage_budget(f32) / R > 1 ? stage_budget(f32) / R : 1);
}

// Radices of the Stockham stages of one direction of a plane transform (16, 8, 4, 3, 2),
// chosen per plane shape and precision so that every stage fits the registers of the plane's
// smallest size class; larger classes run the same stages, so a plane's result does not depend
// on the class it runs in.
struct Plan {
    int n, ns;
    int r[8];
};

struct FFTGeom {
    int nx, ny, nxh, M;          // M = nx/2
    Factors fy, fx;              // radix-4/2/3 factors of ny and M (two-pass transforms)
    Plan py[2], px[2];           // [f32]: stages along y (M columns) and x (ny rows, length M)
    const double2* twx;          // exp(-2 pi i k/nx), k < nx (FP64)
    const double2* twy;          // exp(-2 pi i k/ny), k < ny
    const float2* twx32;         // the same tables rounded to FP32
    const float2* twy32;
    int cls64, cls32;            // plane-size class for FP64 / FP32 compute (-1 = unsupported)
};

// Operands of the plane kernels.  src/dst hold double or float data: bit f of src_f32 marks
// src[f] as float; dst_f32 marks all destinations as float.  Strides are in elements.
// The 2/3 rule of apply_dealias_truncation (DNS_pressure_BC_3D.f90:1163-1193): integer bounds
// kx_max = (nx/2)*2/3, ky_max = (ny/2)*2/3 (:1169-1170); kx index i is kept when i <= kx_max,
// ky index j when j <= ky_max for j <= ny/2 and |j - ny| <= ky_max above (:1176-1190), i.e.
// min(j, ny - j) <= ky_max.
__host__ __device__ inline void dealias_bounds(int nx, int ny, int& kxm, int& kym) {
    kxm = (nx / 2) * 2 / 3;
    kym = (ny / 2) * 2 / 3;
}
// the rule as the plane kernels apply it (FFTArgs::kxm, kym) when on
struct Dealias {
    bool on; int kxm, kym, ny;
    __host__ __device__ __forceinline__ bool keep(int i, int j) const {
        const int jj = j < ny - j ? j : ny - j;
        return !on || (i <= kxm && jj <= kym);
    }
};

struct FFTArgs {
    int nx, ny, nxh, M;
    Plan py, px;
    const void* twx;
    const void* twy;
    const double* xw;
    const double* yw;
    // r2c: the 2/3 dealias rule, on when mask is set: the table (kx index i + nxh * ky index j,
    // 1 = kept) for the two-pass kernels; the plane kernels apply the same rule from its bounds
    // (kept: i <= kxm and min(j, ny - j) <= kym), set by fft_fill
    const unsigned char* mask;
    int kxm, kym;
    const void* src[6];
    void* dst[6];
    long long sstride, dstride;
    unsigned src_f32;
    int dst_f32;
    // zin > 0: item k of a direct launch is plane (k / zin)(zin + 2) + 1 + k % zin, the interior
    // planes of members of zin + 2 planes (the walls skipped); zin = 0: item k is plane k
    int zin;
};

// shape, size classes and stage plans (plans need nx/2 and ny of the form 2^a 3^b)
bool fft_geom(int nx, int ny, FFTGeom& g);
// dynamic LDS bytes of a plane kernel: the plane as M columns x ny rows, the y and x twiddle
// tables, and the c2r side buffers (combined column 0, kx, ky)
size_t plane_lds_bytes(const FFTGeom& g, bool f32);
// fill the geometry fields of a (shape, factors, twiddles of the compute precision)
void fft_fill(const FFTGeom& g, FFTArgs& a, bool f32);
// max complex elements of the half plane for size class c (-1: no such class for f32)
int fft_class_capacity(int c, bool f32);
// the class `levels` steps above class c in order of capacity (more threads per plane; a plane's
// plan fits every such class), or -1 past the largest class of the precision
int fft_class_promote(int c, int levels, bool f32);
// raise the dynamic-LDS limit of the plane kernels of g's size classes (once per process)
hipError_t fft_prepare(const FFTGeom& g);
// grid (nplanes, nfields); f32: compute in FP32 (else FP64)
hipError_t launch_c2r(const FFTGeom& g, int op, FFTArgs a, int nplanes, int nfields, bool f32, hipStream_t s);
hipError_t launch_r2c(const FFTGeom& g, int op, FFTArgs a, int nplanes, int nfields, bool f32, hipStream_t s);
// ragged: every descriptor already filled (fft_fill) for size class cls; lds = dynamic LDS bytes
hipError_t launch_c2r_rag(int cls, int op, bool f32, const FFTArgs* descs, const Tile* map, int ntiles, int nfields,
                          size_t lds, hipStream_t s);
hipError_t launch_r2c_rag(int cls, int op, bool f32, const FFTArgs* descs, const Tile* map, int ntiles, int nfields,
                          size_t lds, hipStream_t s);
// raise the dynamic-LDS limit of every plane-kernel class (once per process)
hipError_t fft_prepare_all();

// Two-pass FP64 transforms of nplanes planes (any plane size with nx/2, ny <= 256): rows and
// columns staged in LDS separately, through scratch (nplanes * 2 (nx/2+1) ny doubles).
// forward: real (stride sstride) -> half-complex (stride dstride), unnormalised; inverse:
// half-complex -> real divided by nx*ny.  Strides in doubles.
struct TwoPassArgs {
    int nx, ny, nxh, M;
    Factors fx, fy;
    const double2* twx;
    const double2* twy;
    const double* src;
    double* dst;
    long long sstride, dstride;
    double scale;
    const unsigned char* mask;
};
hipError_t launch_two_pass(const FFTGeom& g, bool forward, const double* src, double* dst, double* scratch,
                           int nplanes, long long sstride, long long dstride, hipStream_t s);
hipError_t launch_div_spec(const FFTGeom& g, const double* xw, const double* yw, const double* u, const double* v,
                           const double* dw, double* out, int nplanes, hipStream_t s);

// ---- wall-normal contractions and per-mode solves ---------------------------------------
// Batched contexts hold B members of one shape; member b's spectral field starts at
// b * mstride elements and the z-kernels take the member from blockIdx.z.

// device view of a folded operator (setup.h FoldedA; ksp >= ks + 1 k-steps stored per block)
struct FoldOp {
    const double* A;
    const float* Af;             // the same blocks rounded to FP32 (the FP32 contraction), or null
    int rb0, rb1, ks, ksp, rows0, rows1;
    int extra;                   // 1: group-0 row rows0-1 is the dense row after the blocks
    int pb;                      // blocks per group as stored (zero padded); group 1 starts at pb
};
// the dense extra row of a folded operator (valid when extra)
__host__ __device__ inline const double* fold_xrow(const FoldOp& o) {
    return o.A + size_t(2 * o.pb) * o.ksp * 64;
}

// Mode space walked by the Helmholtz and Poisson tiles and by the step's vorticity
// z-derivative.  The fields are transforms of real
// planes, and the step keeps them so: on the kx = 0 and kx = nx/2 columns the modes
// ky > ny/2 are the conjugates of ky < ny/2, and the four self-conjugate modes (0,0),
// (0,ny/2), (nx/2,0), (nx/2,ny/2) are real.  With on (ny even), the tiles walk P = (nx/2) ny
// complex slots, the independent coefficients only.  Slot q is storage mode q wherever that
// mode is independent and not self-conjugate, so a block's slots keep the storage alignment;
// the other slots below P take, in order, two paired slots (the real parts of (0,0) with
// (0,ny/2), and of (nx/2,0) with (nx/2,ny/2)), then the independent modes at or beyond P.  A
// tile stores each mirrored mode as the conjugate of its slot and zeroes the imaginary parts
// of the self-conjugate modes.  At 128 x 128 that is 512 column blocks instead of 520.
// Without on, P = (nx/2+1) ny and slot q is storage mode q.  It is on only where the
// contraction kernels run their four-wave variants (wide: nz of about 98 and more); the one- to
// three-wave variants are compiled without it (their registers would not fit four waves per SIMD).
constexpr int Z_BLOCK_MODES = 16;   // complex slots per column block
struct ZMap {
    int on, nxh, ny, P;
    const int4* tab;             // device: per column of the 2 P, (so, s2o, i, j) (zmap_build)
    const int* bfl;              // device: per column block, ZB_SECOND | ZB_PAIRED | ZB_IDENT
};
// block flags: a column with a second storage column / a paired slot / every column is its own
// storage column with no second one (the kernels then skip the table)
enum { ZB_SECOND = 1, ZB_PAIRED = 2, ZB_IDENT = 4 };
inline ZMap zmap_make(int nx, int ny, bool wide) {
    ZMap z{};
    z.nxh = nx / 2 + 1;
    z.ny = ny;
    z.on = (wide && ny % 2 == 0 && nx >= 4) ? 1 : 0;
    z.P = z.on ? (nx / 2) * ny : z.nxh * ny;
    return z;
}
__host__ __device__ inline int zmap_blocks(const ZMap& z) { return (z.P + Z_BLOCK_MODES - 1) / Z_BLOCK_MODES; }

// Host: the table of the 2 P columns, int4 (so, s2o, i, j) each: so = the column's storage
// real column, s2o = -1 or a second storage column (the mirrored mode's part, which receives
// +v for a real part and -v for an imaginary part, or s2o = so + 1: the zeroed imaginary part
// of a self-conjugate mode), (i, j) = the column's mode; and the flags of the column blocks.
// Returns false if the slots do not come out even (not expected for any shape).
inline bool zmap_build(const ZMap& z, int* tab4, int* bfl) {
    const int W = z.nxh, ny = z.ny, P = z.P, nb = zmap_blocks(z);
    for (int cb = 0; cb < nb; cb++) bfl[cb] = 0;
    auto put = [&](int c, int i, int j, int part, int s2o) {
        const int so = 2 * (i + W * j) + part;
        tab4[4 * c] = so; tab4[4 * c + 1] = s2o; tab4[4 * c + 2] = i; tab4[4 * c + 3] = j;
        if (s2o >= 0) bfl[c / (2 * Z_BLOCK_MODES)] |= ZB_SECOND;
    };
    auto ident_flags = [&]() {
        for (int cb = 0; cb < nb; cb++) {
            bool id = true;
            for (int c = cb * 2 * Z_BLOCK_MODES; c < (cb + 1) * 2 * Z_BLOCK_MODES && c < 2 * P; c++)
                id = id && tab4[4 * c] == c && tab4[4 * c + 1] < 0;
            if (id) bfl[cb] |= ZB_IDENT;
        }
    };
    if (!z.on) {
        for (int q = 0; q < P; q++)
            for (int part = 0; part < 2; part++) put(2 * q + part, q % W, q / W, part, -1);
        ident_flags();
        return true;
    }
    const int h = ny / 2;
    // 0: independent, 1: mirror of (i, ny - j), 2: self-conjugate
    auto kind = [&](int i, int j) {
        const bool edge = i == 0 || i == W - 1;
        return (edge && (j == 0 || j == h)) ? 2 : (edge && j > h) ? 1 : 0;
    };
    auto plain = [&](int q, int i, int j) {
        const bool mir = (i == 0 || i == W - 1) && j > 0 && j < h;
        for (int part = 0; part < 2; part++)
            put(2 * q + part, i, j, part, mir ? 2 * (i + W * (ny - j)) + part : -1);
    };
    int nfree = 0, nhome = 0;
    for (int q = 0; q < P; q++) nfree += kind(q % W, q / W) != 0;
    for (int M = P; M < W * ny; M++) nhome += kind(M % W, M / W) == 0;
    if (nfree != nhome + 2) return false;
    int next = P, paired = 0;   // next storage mode beyond P to place
    for (int q = 0; q < P; q++) {
        const int i = q % W, j = q / W;
        if (kind(i, j) == 0) { plain(q, i, j); continue; }
        if (paired < 2) {
            const int pi = paired == 0 ? 0 : W - 1;
            put(2 * q, pi, 0, 0, 2 * pi + 1);
            put(2 * q + 1, pi, h, 0, 2 * (pi + W * h) + 1);
            bfl[q / Z_BLOCK_MODES] |= ZB_PAIRED;
            paired++;
            continue;
        }
        while (kind(next % W, next / W) != 0) next++;
        plain(q, next % W, next / W);
        next++;
    }
    ident_flags();
    return true;
}
struct ZCol { int so, s2o, i, j; };
// device: column col from the table, or (ident: its block has ZB_IDENT) from its index
__device__ __forceinline__ ZCol zcol_t(const ZMap& z, int col, bool ident) {
    if (ident) {
        const int q = col >> 1, j = q / z.nxh;
        return ZCol{col, -1, q - j * z.nxh, j};
    }
    const int4 t = gmem(z.tab)[col];
    return ZCol{t.x, t.y, t.z, t.w};
}

struct ZArgs {
    long long mstride;
    int nz, H2, nxh, ny, nx;     // H2 = 2 * (nx/2+1) * ny real columns per plane
    FoldOp d1;                   // folded D1: s (centre group) on X_O | r on X_E
    const double* in[3];
    double* out[3];              // float when out_f32
    const double* xw; const double* yw;
    int out_f32;
    int d1_f32;                  // the vorticity's D1 on FP32 MFMA (d1.Af), FP32 accumulation
    int zwaves;                  // least waves per column block (z_variant_widen; 0 = natural)
    // packed: the columns walked are the per-mode solves' slots (ZMap; the vorticity of the step),
    // else the storage columns
    int packed;
    ZMap zm;
};
enum ZMode { Z_PLAIN = 0, Z_VORT = 1 };
// row-block pairs of a folded operator
inline int fold_nrb(const FoldOp& o) { return o.rb0 > o.rb1 ? o.rb0 : o.rb1; }
// Kernel variant of a contraction kernel whose operators have at most nrb row-block pairs:
// 0, 1, 2 = one wave per pair (1 to 3 pairs; results unfold through LDS so that every thread
// stores), 3 = four waves with one pair each, 4 = four waves with two pairs each (the stored
// operators are zero padded to 4 ceil(nrb / 4) blocks per group); -1 = more than 8 pairs.
inline int z_variant(int nrb) { return nrb < 4 ? nrb - 1 : nrb <= 8 ? 2 + (nrb + 3) / 4 : -1; }
inline bool z_variant_lx(int var) { return var >= 0 && var < 3; }
// the variant of a kernel applying the operators o[0..n): -1 unless all are padded alike
inline int z_variant_of(const FoldOp* const* o, int n) {
    int m = 0;
    for (int i = 0; i < n; i++) m = fold_nrb(*o[i]) > m ? fold_nrb(*o[i]) : m;
    for (int i = 0; i < n; i++)
        if (o[i]->pb != 4 * ((m + 3) / 4)) return -1;
    return z_variant(m);
}
// The variant of a column block given the least waves per block asked for (zwaves, a tile
// shape an ensemble's tuner picks; 0 = one wave per row-block pair).  One- and two-wave
// variants widen to two or four waves (one row-block pair per wave, the extra waves idle in the
// contractions); the results do not change (contract.h XAcc: the extra row is summed by row
// class whatever the width).  Three waves keep their width.
inline int z_variant_widen(int var, int zwaves) {
    if (var < 0 || var > 1 || zwaves <= var + 1) return var;
    return zwaves >= 4 ? 3 : 1;
}
inline int zd_variant(const ZArgs& a) {
    const FoldOp* o[1] = {&a.d1};
    return z_variant_widen(z_variant_of(o, 1), a.zwaves);
}
hipError_t launch_zderiv(int mode, const ZArgs& a, int nf, int nmembers, hipStream_t s);
size_t zderiv_lds_bytes(const ZArgs& a);

struct FlowState;
struct FlowMember;
struct FlowArgs {
    int nz;
    double dt, nxny, wsum;
    const double* uh;            // u-hat: reads Re u(0,0,k) at stride H2
    long long H2, mstride;
    const double* wz;
    const FlowMember* mem;
    FlowState* st;
};

struct Mode0Args {
    long long mstride;
    int nz, H2;
    const double* e0;            // row-major nz x nz: phi(0,0,:) = E0 w*(0,0,:)
    const double* f0;            // row-major nz x nz: D1 phi(0,0,:) = F0 w*(0,0,:)
    double dt;
    int fused;
    const double* us; const double* vs; const double* ws;
    double* u; double* v; double* w; double* p;
    double* phi; double* dphi;
};
struct HelmArgs {
    long long mstride, fstride;  // member strides of the fields and of hF.A (per-member Re)
    int nz, m, H2, nxh, ny, nx;
    int zwaves;                  // least waves per column block (z_variant_widen; 0 = natural)
    FoldOp hF;                   // [T | f T D2_I] folded: even | odd eigen rows (member 0's A)
    FoldOp hB;                   // p | q
    const double* c1; const double* c2; const double* hl;   // m per member, even rows first
    const double* fvec;          // theta dt / Re per member
    double dt, nxny;
    int cn;
    const double* xw; const double* yw;
    const double* uh[3]; const double* sh[3]; double* out[3];   // sh float when s_f32
    const double* ph;
    int s_f32;
    const FlowState* flow;       // per member: current pressure gradient G
    // flow_fused: the block of column block 0, component 0 runs the flow controller (fl) for its
    // member before reading G (the fused step; the stage API launches it separately)
    FlowArgs fl;
    int flow_fused;
    ZMap zm;                     // the tiles' mode space
};
hipError_t launch_helmholtz(const HelmArgs& a, int nmembers, hipStream_t s);
inline int helm_variant(const HelmArgs& a) {
    const FoldOp* o[2] = {&a.hF, &a.hB};
    return z_variant_widen(z_variant_of(o, 2), a.zwaves);
}
// dynamic LDS bytes of the Helmholtz / Poisson kernels (staging, then the two z panels)
size_t helm_lds_bytes(const HelmArgs& a);

struct PoisArgs {
    long long mstride;
    int nz, H2, nxh, ny, nx;
    int zwaves;                  // least waves per column block (z_variant_widen; 0 = natural)
    FoldOp pF, pB, pDB;          // folded [Tf | Tf D1], Bf, D1 Bf
    const double* pl;            // even eigenvalues first
    double dt;
    int fused;                   // 1: apply correction + walls; 0: write phi, D1 phi only
    int weak;                    // weak-form RHS: div = -(i xw u* + i yw v*) on the raw wavenumbers
    const double* xw; const double* yw;
    const double* us; const double* vs; const double* ws;   // u*, v*, w* spectral
    double* u; double* v; double* w; double* p;             // fused outputs
    double* phi; double* dphi;                              // stage outputs
    // mode0: wave 0 of the block of column block cb also solves mode (0,0) at the nodes
    // k = cb, cb + ncb, ... of its member (m0; the Poisson tiles skip that mode)
    Mode0Args m0;
    int mode0;
    ZMap zm;                     // the tiles' mode space
};
hipError_t launch_poisson(const PoisArgs& a, int nmembers, hipStream_t s);
// device facts the z-kernel launchers use (compute units); once per process, before any launch
hipError_t z_prepare();
inline int pois_variant(const PoisArgs& a) {
    const FoldOp* o[3] = {&a.pF, &a.pB, &a.pDB};
    return z_variant_widen(z_variant_of(o, 3), a.zwaves);
}
size_t pois_lds_bytes(const PoisArgs& a);


struct CorrArgs {
    int nz, H2, nxh, ny, nx;
    const double* xw; const double* yw;
    const double* us; const double* vs; const double* ws;
    const double* phi; const double* dphi;
    double* u; double* v; double* w; double* p;
    double dt;
};
hipError_t launch_correction(const CorrArgs& a, hipStream_t s);
hipError_t launch_zero_walls(int nz, int H2, long long mstride, int nmembers, double* u, double* v, double* w,
                             hipStream_t s);

// One stage of the 4-stage 1/irk convection (runge_kutta_convection, DNS_pressure_BC_2D.f90:
// 931-1014) on the spectral state: out_c = un_c + a S_c for the three components (a = dt / irk),
// zero on the wall planes (apply_wall_bc after each stage)
struct RKStageArgs {
    long long n;                 // doubles per field over all members (members * nz * H2)
    long long H2; int nz;
    const double* un[3]; const void* s[3]; int s_f32;
    double* out[3];
    double a;
};
hipError_t launch_rk_stage(const RKStageArgs& a, hipStream_t s);

// jacgs (DNS_pressure_BC_2D.f90:2361-2476) on each real column (a mode's real or imaginary part)
// of the 3D Poisson system of solve_pressure_3d_standard (DNS_pressure_BC_3D.f90:1680-1711):
// A = -St - diag((kx^2 + ky^2) w (ybar/2)), mode (0,0) with row and column 1 replaced by the
// identity; b = (div/dt) (ybar/2) w with div = i kx_eff u* + i ky_eff v* + D1 w*; x0 = 0; Jacobi
// preconditioner; stop when ||r||_2 <= tol or after niter iterations.  One wavefront per column.
struct JacgsArgs {
    long long mstride;
    int nz, H2, nxh, nx, ny, niter;
    const double* st;            // nz x nz, symmetric: (2/ybar) sum_m D1(m,k) D1(m,n) w_m
    const double* wz;
    const double* xw; const double* yw;
    const double* us; const double* vs; const double* dws;   // u*, v*, D1 w* (spectral)
    double* phi;
    double dt, ybar, tol;
    unsigned long long* iters;   // total iterations (atomic)
    int* nonconv;                // systems that met niter without converging (atomic)
};
hipError_t launch_jacgs(const JacgsArgs& a, int nmembers, hipStream_t s);
// zero the two counters of launch_jacgs (a kernel: no memset node in captured graphs)
hipError_t launch_jacgs_clear(unsigned long long* iters, int* nonconv, hipStream_t s);

// PI / constant forcing state, device resident (update_pressure_gradient_3d), one per member.
// zero_wallgrad: the Helmholtz wall values of the next step use a zero pressure gradient (the
// reference's grad_p_prev after initialize_fields_3d or a restart read, which is set only by
// pressure_correction_3d, DNS_pressure_BC_3D.f90:957,1923); 0 once a pressure correction has
// run, when the wall gradient is that of the resident p.
struct FlowState {
    double G, Ub, err, err_int, err_prev, gsum, gavg;
    int gcount, zero_wallgrad;
};
struct FlowMember {
    double target_G, target_Ub, gain, re;
    int method, freq, istart, pad;
};
hipError_t launch_flow_control(const FlowArgs& a, int nmembers, hipStream_t s);
// zero_wallgrad = 0 for members [0, nmembers) (a small kernel: no memset node in captured graphs)
hipError_t launch_clear_wallgrad(FlowState* st, int nmembers, hipStream_t s);

// ragged variants of the step kernels (var = zd/helm/pois_variant of the groups; lds = dynamic
// LDS bytes, the largest over the launch's groups)
hipError_t launch_zderiv_rag(int var, int mode, bool f32, const ZArgs* d, const Tile* map, int ntiles, int nf, size_t lds,
                             hipStream_t s);
hipError_t launch_helmholtz_rag(int var, bool sf32, const HelmArgs* d, const Tile* map, int ntiles, size_t lds,
                                hipStream_t s);
hipError_t launch_poisson_rag(int var, const PoisArgs* d, const Tile* map, int ntiles, size_t lds, hipStream_t s);

// physical initial fields of one member from the separable form (setup.h PlanarInit):
// u = U[k] + scale(u2 wd[k]), v = scale(v2 wd[k]), w = 0, walls zero
struct InitArgs {
    long long npl; int nz;
    const double* u2; const double* v2; const double* U; const double* wd;
    double sf; int scaled, perturbed;
    double* u; double* v; double* w;
};
hipError_t launch_init_fields(const InitArgs& a, hipStream_t s);
// y = x for n doubles (n even)
hipError_t launch_copy(const double* x, double* y, size_t n, hipStream_t s);

// reductions over physical fields: max|u|,|v|,|w|, sum 0.5|u|^2 w_k, max|d| (any may be null)
struct ReduceArgs {
    long long npl; int nz;
    const double* u; const double* v; const double* w; const double* d;
    const double* wz;
    double* partial;             // nblocks * 5
    double* result;              // 5
    int nblocks;
};
hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s);

// 4096 one-hot products of v_mfma_f64_16x16x4f64 (out: 4096 * 256 doubles)
hipError_t launch_mfma_probe(double* out, hipStream_t s);
hipError_t launch_mfma_probe_f32(double* out, hipStream_t s);

}  // namespace dns
