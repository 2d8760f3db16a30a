// context.h -- the resident context behind dns_ctx (include/f90dns.h) and the argument
// blocks of its step kernels, shared by the single-context path and the ragged ensemble.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/f90dns.h"
#include "kernels.h"
#include "setup.h"

struct dns_ctx {
    dns_params p;                 // member 0 (the shape fields are equal for every member)
    std::vector<dns_params> mp;   // every member
    int B = 1;
    int nx, ny, nz, nxh, H2, npl;
    size_t N, NS;                 // per member: physical doubles, spectral doubles
    double theta, ybar;
    int rb, rbi;
    bool mixed = false;           // FP32 transforms, product and omega/S storage; FP64 solves
    bool d1f32 = false;           // and the vorticity's D1 contraction on FP32 MFMA (D1sw32)
    hipStream_t stream = nullptr;
    dns::FFTGeom g;
    dns::Operators op;                 // member 0; the Re-independent parts are shared
    size_t hf_stride = 0;         // doubles per member of the stacked Helmholtz forward operator
    // constants
    float *twx32 = nullptr, *twy32 = nullptr;
    double *twx = nullptr, *twy = nullptr, *xw = nullptr, *yw = nullptr, *wz = nullptr, *e0 = nullptr, *f0 = nullptr;
    unsigned char* mask = nullptr;
    float* D1sw32 = nullptr;
    double *D1sw = nullptr, *hF = nullptr, *hB = nullptr, *hc1 = nullptr, *hc2 = nullptr, *hl = nullptr, *fvec = nullptr;
    double *pF = nullptr, *pB = nullptr, *pDB = nullptr, *pl = nullptr;
    int4* ztab = nullptr;         // mode space of the per-mode solves: per column (zmap_build)
    int* zbfl = nullptr;          //   and per column block
    dns::ZMap zm{};               //   its shape (tab, bfl set)
    // fields, B members each
    double *uh = nullptr, *vh = nullptr, *wh = nullptr, *ph = nullptr;
    double *us = nullptr, *vs = nullptr, *ws = nullptr;
    double *ox = nullptr, *oy = nullptr, *sh[3] = {nullptr, nullptr, nullptr};
    double *phys[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double *phi = nullptr, *dphi = nullptr, *tmp = nullptr;
    double* io = nullptr;         // one member's spectral scratch for the two-pass FP64 transforms
    dns::FlowState* flow = nullptr;
    dns::FlowMember* fmem = nullptr;
    // members whose device FlowState::zero_wallgrad is set (fields set with p): cleared on the
    // stream after the next pressure correction
    std::vector<char> zwg;
    double *red_partial = nullptr, *red_result = nullptr;
    // pressure_solver 1 (jacgs): the symmetric stiffness matrix and the solve counters
    double* stiff = nullptr;
    unsigned long long* jac_iters = nullptr;
    int* jac_nonconv = nullptr;
    int red_blocks = 512;
    std::vector<void*> allocs;
    size_t bytes = 0;
    double time = 0.0;
    int step = 0;
    std::string err;
    // timing
    std::vector<hipEvent_t> ev;
    std::vector<double> step_ms;
    // graph of one step
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace dns {
// the contraction kernels of these operators all run their four-wave variants (the per-mode
// solves' packed mode space is on only then)
bool ops_wide(const Operators& o);
// Argument blocks of the fused step's kernels over all members of c (plane-kernel blocks
// without their geometry: fft_fill adds it).
FFTArgs ctx_c2r_nl_args(dns_ctx* c);
FFTArgs ctx_r2c_cross_args(dns_ctx* c);
ZArgs ctx_vort_args(dns_ctx* c);
FlowArgs ctx_flow_args(dns_ctx* c);
HelmArgs ctx_helm_args(dns_ctx* c);
PoisArgs ctx_pois_args(dns_ctx* c, int fused);
Mode0Args ctx_mode0_args(dns_ctx* c, int fused);
// after a step's pressure correction is enqueued on s: clears the zero-wall-gradient state of
// members whose fields were set with p (no-op otherwise)
int ctx_after_correction(dns_ctx* c, hipStream_t s);
}  // namespace dns

void dns_set_create_error(const std::string& msg);
bool dns_check_gfx950(std::string& why);
bool dns_make_factors(int n, dns::Factors& f);
