// lsrtm2d.hip - 2D elastic / viscoelastic context behind include/erm2d.h.
//
// The reference's 2D directory has the kernels but no host driver; this file
// is the driver: allocation, model setup (cuda_cal_viscoelastic,
// viscoelastic_equation.cu:22-58), the modeling / RTM time loops, and the
// LSRTM operators (Born = background and Born states stepped together;
// adjoint = background pass storing the strain increments and stress forces
// of the model region for every step, then the exact transpose stepped back).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <string>
#include <thread>
#include <vector>

#include "../../include/erm2d.h"
#include "common.hpp"
#include "kernels2d.hpp"

using namespace erm2;

namespace {
thread_local std::string g_err2;

// Staging of erm2d_angle_gathers_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel or a copy (on the host it is a stand-in, not a
// device address): erm2d_create resolves each device's own address once,
// with that device current, and the entry point uses that address.
constexpr size_t kScratchFloats = (size_t)1 << 23;  // 32 MiB
static __device__ __align__(256) float scratch[kScratchFloats];
constexpr int kMaxDevices = 64;
float* g_scratch_ptr[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

// the calling thread's current device must be `device`
int resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return ERM_E_PARAM;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return ERM_E_DEVICE;
    g_scratch_ptr[device] = static_cast<float*>(ptr);
    return 0;
}
const double kC6[7] = {0.0, 1.2213363647, -0.096931457519, 0.017447662353,
                       -0.0029672895159, 0.00035900539822, -0.00002184781161};
const double kC4[5] = {0.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0};
constexpr int TXR = 4;       // rows per workgroup (64 x 4 = 256 threads)
constexpr int TXR_TALL = 8;  // the taller stencil workgroups of large launches (erm2d_params.tile_rows)

struct Sls {
    double fM, fU, gP[3], gS[3], taus[3];
};
// cuda_cal_viscoelastic (viscoelastic_equation.cu:43-56) for one mechanism;
// tau-method (Blanch et al. 1995) for three: tau_sig_l = 1/(2 pi f_l),
// f_l = freq {1/4, 1, 4}, tau_eps_l = tau_sig_l (1 + tau), Q(2 pi freq) = Q.
Sls sls_params(int L, double freq, double qp, double qs) {
    Sls s{};
    const double w = 2 * 3.1415926 * freq;
    if (L == 0) { s.fM = s.fU = 1.0; return s; }
    if (L == 1) {
        double tao = 1.0 / w * (std::sqrt(1.0 + 1.0 / qp / qp) - 1.0 / qp);
        double sp = 1.0 / (w * w * tao);
        double ss = (1 + w * tao * qs) / (w * qs - w * w * tao);
        s.taus[0] = tao;
        s.fM = sp / tao;
        s.fU = ss / tao;
        s.gP[0] = (1.0 - sp / tao) / tao;
        s.gS[0] = (1.0 - ss / tao) / tao;
        return s;
    }
    const double f3[3] = {freq / 4, freq, freq * 4};
    double S1 = 0, S2 = 0;
    for (int l = 0; l < 3; ++l) {
        s.taus[l] = 1.0 / (2 * M_PI * f3[l]);
        double x = w * s.taus[l];
        S1 += x / (1 + x * x);
        S2 += x * x / (1 + x * x);
    }
    double tP = 1.0 / (qp * S1 - S2), tS = 1.0 / (qs * S1 - S2);
    s.fM = 1 + 3 * tP;
    s.fU = 1 + 3 * tS;
    for (int l = 0; l < 3; ++l) { s.gP[l] = -tP / s.taus[l]; s.gS[l] = -tS / s.taus[l]; }
    return s;
}
}  // namespace

struct erm2d_ctx {
    erm2d_params p;
    int device = 0;
    hipStream_t stream = nullptr;
    int R = 4, L = 0, nnx = 0, nnz = 0, pz = 0;
    size_t ncell = 0, mn = 0;
    Coef2 coef{};
    float cx = 0, cz = 0, dtr = 0;
    float* fA[8] = {nullptr};
    float* rA[3][4] = {{nullptr}};
    float* fB[8] = {nullptr};
    float* rB[3][4] = {{nullptr}};
    float* W[3] = {nullptr};
    float* dv = nullptr;
    float* cu = nullptr;
    struct Mod { float *M = nullptr, *U = nullptr, *B = nullptr, *kP[3] = {0}, *kS[3] = {0}, *decay[3] = {0};
                 float *vp = nullptr, *vs = nullptr; bool have = false; } mod[2];
    int last_model = 1;     // model of the last pass (decom_new scaling)
    bool cpml = false;      // C-PML boundary (modeling and RTM)
    float* psi = nullptr;   // 10 memory variables x ncell
    float* pcx = nullptr;   // [4][nnx] (a_full, b_full, a_half, b_half)
    float* pcz = nullptr;   // [4][nnz]
    double vmax_loaded = 0;
    float* att = nullptr;
    std::vector<float> wavelet;
    long src = -1;
    long* rec_idx = nullptr;
    float* rec = nullptr;   // [2][lt][nrec]
    float* rec2 = nullptr;  // second record buffer (residual / Born data)
    int* ex_time = nullptr;
    float* ex[6] = {nullptr};  // amp tp vxp vzp, receiver vxp vzp at the imaging step
    int* ex_tp_time = nullptr; // tp excitation (caculate_ex_tp_time_new)
    float* ex_tp_amp = nullptr;
    float* img[3] = {nullptr};
    float* hist = nullptr;  // [lt][5][mn]: E (3) then F (2); [5][mn] (one step) with adjoint_source strips
    float* ls_strip = nullptr;  // adjoint_source strips: [lt][2][4][nr] (stress side, velocity side)
    float* dm = nullptr;    // [3][mn]
    float* grad = nullptr;  // [3][mn]
    float* illum = nullptr; // source illumination accumulator (padded grid)
    hipEvent_t ev[2] = {nullptr, nullptr};
    double last_ms = 0;
    float qsign = 1.f;                      // -1 during a Q-compensated receiver pass
    float dsgn = 1.f;                       // -1 during a reversed source step (erm2d_rtm_xcorr)
    float* psiB = nullptr;                  // C-PML memory variables of state B (erm2d_rtm_xcorr)
    // erm2d_rtm_xcorr buffers (allocated on first use)
    float* xc_img = nullptr;                // [5][mn]
    int* xc_vote = nullptr;                 // per cell, padded grid
    int* xc_sign = nullptr;
    float* xc_hist = nullptr;               // strategy 0: [lt_forward][3][mn]
    size_t xc_hist_n = 0;
    long* xc_ring = nullptr;                // strategy 2: ring cell indices
    long xc_nr = 0;
    float* xc_strip = nullptr;              // strategy 2: [lt_forward][2][4][nr] (stress side, velocity side)
    size_t xc_strip_n = 0;
    std::vector<float> bg_host[5];          // background model as loaded (vp vs rho qp qs), for workers
    std::vector<erm2d_ctx*> workers;        // extra contexts (own streams) for concurrent shots
    // Batched LSRTM operators (erm2d_set_batch): per-shot copies of the two
    // states, the adjoint scratch, the E/F history, the gradients and the
    // records for up to cap shots, each array [cap][...]; nb > 0 while a
    // batched operator runs (the launches then cover nb shots along z).
    struct Batch {
        int cap = 0;
        float *A[8] = {nullptr}, *B[8] = {nullptr}, *rA[3][4] = {{nullptr}}, *rB[3][4] = {{nullptr}}, *W[3] = {nullptr};
        float *hist = nullptr, *grad = nullptr, *rec = nullptr;
        float* strip = nullptr;  // adjoint_source strips: [cap][lt][2][4][nr]
        long* srcs = nullptr;
    } bt;
    int nb = 0;
    erm2d_cg_monitor_fn monitor = nullptr;
    void* monitor_user = nullptr;
    std::string err;
    int fail(int code, const char* fmt, ...) {
        va_list ap;
        va_start(ap, fmt);
        err = erm::vformat(fmt, ap);
        va_end(ap);
        return code;
    }
};

#define HC(c, call)                                                                                      \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return (c)->fail(ERM_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_));     \
    } while (0)

namespace {

template <typename T>
int dmal(erm2d_ctx* c, T** p, size_t n) {
    if (n == 0) { *p = nullptr; return 0; }
    if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) return c->fail(ERM_E_NOMEM, "hipMalloc(%zu) failed", n * sizeof(T));
    return 0;
}
int fillf(erm2d_ctx* c, float* p, size_t n, float v) {
    if (!p) return 0;
    hipLaunchKernelGGL(k2_fill, dim3(512), dim3(256), 0, c->stream, p, (long)n, v);
    HC(c, hipGetLastError());
    return 0;
}

// exact-transpose adjoint from boundary strips (adjoint_source 2): the E/F
// history holds one step per shot, and the ring strips every step
bool ls_strips(const erm2d_ctx* c) { return c->p.adjoint_source == ERM2D_SRC_STRIPS; }
int hist_steps(const erm2d_ctx* c) { return ls_strips(c) ? 1 : c->p.lt; }
std::vector<long> xcorr_ring(const erm2d_ctx* c);
// the ring cell indices of the strips strategies (built on first use)
int ensure_ring(erm2d_ctx* c);

G2 geo(const erm2d_ctx* c) {
    G2 g;
    g.nnx = c->nnx; g.nnz = c->nnz; g.pz = c->pz;
    g.cx = c->cx; g.cz = c->cz; g.dtr = c->dtr;
    g.att = c->att;
    g.bl = c->p.bl; g.bu = c->p.bu; g.nx = c->p.nx; g.nz = c->p.nz;
    g.qsign = c->qsign;
    g.dsgn = c->dsgn;
    g.psi = c->cpml ? c->psi : nullptr;
    g.pcx = c->pcx;
    g.pcz = c->pcz;
    g.ncell = (long)c->ncell;
    g.sst = c->nb ? (long)c->ncell : 0;
    g.hst = c->nb ? (long)hist_steps(c) * 5 * (long)c->mn : 0;
    g.gst = c->nb ? 3 * (long)c->mn : 0;
    g.srcs = c->nb ? c->bt.srcs : nullptr;
    return g;
}
F2 state(erm2d_ctx* c, bool B) {
    F2 f;
    float** a = c->nb ? (B ? c->bt.B : c->bt.A) : (B ? c->fB : c->fA);
    f.vx = a[0]; f.vz = a[1]; f.xx = a[2]; f.zz = a[3]; f.xz = a[4]; f.tp = a[5]; f.vxp = a[6]; f.vzp = a[7];
    for (int l = 0; l < 3; ++l)
        for (int q = 0; q < 4; ++q)
            f.r[l][q] = c->nb ? (B ? c->bt.rB[l][q] : c->bt.rA[l][q]) : (B ? c->rB[l][q] : c->rA[l][q]);
    return f;
}
// the history / scratch, gradient, record and adjoint-scratch buffers of the
// running operator (batched: shot 0's; the kernels offset the others)
float* hist_of(erm2d_ctx* c) { return c->nb ? c->bt.hist : c->hist; }
float* grad_of(erm2d_ctx* c) { return c->nb ? c->bt.grad : c->grad; }
float* rec_of(erm2d_ctx* c) { return c->nb ? c->bt.rec : c->rec2; }
float* const* w_of(erm2d_ctx* c) { return c->nb ? c->bt.W : c->W; }
int nshot_of(const erm2d_ctx* c) { return c->nb ? c->nb : 1; }
M2 model(const erm2d_ctx* c, int which) {
    M2 m;
    const auto& d = c->mod[which];
    m.M = d.M; m.U = d.U; m.B = d.B;
    m.vp = d.vp; m.vs = d.vs;
    for (int l = 0; l < 3; ++l) { m.kP[l] = d.kP[l]; m.kS[l] = d.kS[l]; m.decay[l] = d.decay[l]; }
    return m;
}
Ex2 excite(const erm2d_ctx* c) {
    Ex2 e;
    e.time = c->ex_time; e.amp = c->ex[0]; e.tp = c->ex[1]; e.vxp = c->ex[2]; e.vzp = c->ex[3];
    e.rvxp = c->ex[4]; e.rvzp = c->ex[5];
    e.img_tp = c->img[0]; e.img_pp = c->img[1]; e.img_ps = c->img[2];
    e.tp_time = c->ex_tp_time; e.tp_amp = c->ex_tp_amp; e.tp_shift = c->p.ex_tp_shift;
    return e;
}
dim3 grid2(const erm2d_ctx* c, int rows = TXR) { return dim3((c->nnz + TZ - 1) / TZ, (c->nnx + rows - 1) / rows, nshot_of(c)); }
// x rows per workgroup of the stress / velocity / adjoint stencil launches
// (erm2d_params.tile_rows): the taller workgroups re-read fewer x-halo rows
// per row but halve the workgroup count, so by default only launches that
// still have >= 4096 of them (batched shots) use them
int rows(const erm2d_ctx* c) {
    if (c->p.tile_rows) return c->p.tile_rows;
    const long wg = (long)((c->nnz + TZ - 1) / TZ) * ((c->nnx + TXR_TALL - 1) / TXR_TALL) * nshot_of(c);
    return wg >= 4096 ? TXR_TALL : TXR;
}
dim3 grid_pt(const erm2d_ctx* c) { return dim3((c->nnz + 255) / 256, c->nnx, nshot_of(c)); }

// geometry of a launch on state A (the C-PML memory variables belong to it) or B
G2 geo_of(const erm2d_ctx* c, bool B) {
    G2 g = geo(c);
    if (B) g.psi = c->cpml && !c->nb ? c->psiB : nullptr;
    return g;
}
template <int R, int L, bool SAVE>
void stress_r(erm2d_ctx* c, bool B, int which, long src, float sv, float* Eh) {
    if (rows(c) == TXR_TALL)
        hipLaunchKernelGGL((k2_stress<R, L, TXR_TALL, SAVE>), grid2(c, TXR_TALL), dim3(TZ, TXR_TALL), 0, c->stream,
                           state(c, B), model(c, which), geo_of(c, B), c->coef, src, sv, Eh);
    else
        hipLaunchKernelGGL((k2_stress<R, L, TXR, SAVE>), grid2(c), dim3(TZ, TXR), 0, c->stream, state(c, B),
                           model(c, which), geo_of(c, B), c->coef, src, sv, Eh);
}
template <int R, bool SAVE>
void stress_l(erm2d_ctx* c, bool B, int which, long src, float sv, float* Eh) {
    if (c->L == 0) stress_r<R, 0, SAVE>(c, B, which, src, sv, Eh);
    else if (c->L == 1) stress_r<R, 1, SAVE>(c, B, which, src, sv, Eh);
    else stress_r<R, 3, SAVE>(c, B, which, src, sv, Eh);
}
void do_stress(erm2d_ctx* c, bool B, int which, long src, float sv, float* Eh) {
    if (c->R == 4) { if (Eh) stress_l<4, true>(c, B, which, src, sv, Eh); else stress_l<4, false>(c, B, which, src, sv, Eh); }
    else { if (Eh) stress_l<6, true>(c, B, which, src, sv, Eh); else stress_l<6, false>(c, B, which, src, sv, Eh); }
}
template <int R, int MODE, bool SAVE>
void vel_r(erm2d_ctx* c, bool B, int which, int it, float* Fh) {
    if (rows(c) == TXR_TALL)
        hipLaunchKernelGGL((k2_velocity<R, TXR_TALL, MODE, SAVE>), grid2(c, TXR_TALL), dim3(TZ, TXR_TALL), 0, c->stream,
                           state(c, B), model(c, which), geo_of(c, B), c->coef, excite(c), it, Fh);
    else
        hipLaunchKernelGGL((k2_velocity<R, TXR, MODE, SAVE>), grid2(c), dim3(TZ, TXR), 0, c->stream, state(c, B),
                           model(c, which), geo_of(c, B), c->coef, excite(c), it, Fh);
}
template <int R>
void vel_m(erm2d_ctx* c, bool B, int which, int mode, int it, float* Fh) {
    if (Fh) vel_r<R, 0, true>(c, B, which, it, Fh);
    else if (mode == 1) vel_r<R, 1, false>(c, B, which, it, nullptr);
    else if (mode == 2) vel_r<R, 2, false>(c, B, which, it, nullptr);
    else vel_r<R, 0, false>(c, B, which, it, nullptr);
}
void do_velocity(erm2d_ctx* c, bool B, int which, int mode, int it, float* Fh) {
    if (c->R == 4) vel_m<4>(c, B, which, mode, it, Fh);
    else vel_m<6>(c, B, which, mode, it, Fh);
    if (c->p.ex_vzp_avg && !Fh && (mode == 1 || mode == 2)) {
        if (mode == 1) hipLaunchKernelGGL(k2_ex_avg<1>, grid_pt(c), dim3(256), 0, c->stream, state(c, B), excite(c), geo(c), it, c->R);
        else hipLaunchKernelGGL(k2_ex_avg<2>, grid_pt(c), dim3(256), 0, c->stream, state(c, B), excite(c), geo(c), it, c->R);
    }
}
int para(const erm2d_ctx* c) { return c->p.inversion_para == 1 || c->p.inversion_para == 3 ? c->p.inversion_para : 2; }
void born_stress_src(erm2d_ctx* c, const float* Eh) {
    dim3 g = grid_pt(c);
    if (c->L == 0) hipLaunchKernelGGL((k2_born_stress_src<0>), g, dim3(256), 0, c->stream, state(c, true), model(c, 1), geo(c), c->dm, Eh, para(c));
    else if (c->L == 1) hipLaunchKernelGGL((k2_born_stress_src<1>), g, dim3(256), 0, c->stream, state(c, true), model(c, 1), geo(c), c->dm, Eh, para(c));
    else hipLaunchKernelGGL((k2_born_stress_src<3>), g, dim3(256), 0, c->stream, state(c, true), model(c, 1), geo(c), c->dm, Eh, para(c));
}
template <int R, int TR>
void adj_v_r(erm2d_ctx* c, const float* Fh) {
    hipLaunchKernelGGL((k2a_v<R, TR>), grid2(c, TR), dim3(TZ, TR), 0, c->stream, state(c, true), model(c, 1), geo(c),
                       c->coef, Fh, grad_of(c));
}
void adj_v(erm2d_ctx* c, const float* Fh) {
    const bool t = rows(c) == TXR_TALL;
    if (c->R == 4) { if (t) adj_v_r<4, TXR_TALL>(c, Fh); else adj_v_r<4, TXR>(c, Fh); }
    else { if (t) adj_v_r<6, TXR_TALL>(c, Fh); else adj_v_r<6, TXR>(c, Fh); }
}
void adj_s1(erm2d_ctx* c, const float* Eh) {
    dim3 g = grid_pt(c);
    float* gr = grad_of(c);
    float* const* W = w_of(c);
    if (c->L == 0) hipLaunchKernelGGL((k2a_s1<0>), g, dim3(256), 0, c->stream, state(c, true), model(c, 1), geo(c), Eh, gr, W[0], W[1], W[2], c->R);
    else if (c->L == 1) hipLaunchKernelGGL((k2a_s1<1>), g, dim3(256), 0, c->stream, state(c, true), model(c, 1), geo(c), Eh, gr, W[0], W[1], W[2], c->R);
    else hipLaunchKernelGGL((k2a_s1<3>), g, dim3(256), 0, c->stream, state(c, true), model(c, 1), geo(c), Eh, gr, W[0], W[1], W[2], c->R);
}
template <int R, int TR>
void adj_s2_r(erm2d_ctx* c) {
    float* const* W = w_of(c);
    hipLaunchKernelGGL((k2a_s2<R, TR>), grid2(c, TR), dim3(TZ, TR), 0, c->stream, state(c, true), geo(c), c->coef, W[0],
                       W[1], W[2]);
}
void adj_s2(erm2d_ctx* c) {
    const bool t = rows(c) == TXR_TALL;
    if (c->R == 4) { if (t) adj_s2_r<4, TXR_TALL>(c); else adj_s2_r<4, TXR>(c); }
    else { if (t) adj_s2_r<6, TXR_TALL>(c); else adj_s2_r<6, TXR>(c); }
}

// reference-mode adjoint steps on state B (see k2r_s / k2r_v)
template <int R, int L, int TR>
void ref_step_rlt(erm2d_ctx* c, const float* Eh) {
    hipLaunchKernelGGL((k2r_s<R, L, TR>), grid2(c, TR), dim3(TZ, TR), 0, c->stream, state(c, true), model(c, 1), geo(c),
                       c->coef);
    hipLaunchKernelGGL((k2r_v<R, L, TR>), grid2(c, TR), dim3(TZ, TR), 0, c->stream, state(c, true), model(c, 1), geo(c),
                       c->coef, Eh, grad_of(c));
}
template <int R, int L>
void ref_step_rl(erm2d_ctx* c, const float* Eh) {
    if (rows(c) == TXR_TALL) ref_step_rlt<R, L, TXR_TALL>(c, Eh);
    else ref_step_rlt<R, L, TXR>(c, Eh);
}
template <int R>
void ref_step_r(erm2d_ctx* c, const float* Eh) {
    if (c->L == 0) ref_step_rl<R, 0>(c, Eh);
    else if (c->L == 1) ref_step_rl<R, 1>(c, Eh);
    else ref_step_rl<R, 3>(c, Eh);
}
void ref_step(erm2d_ctx* c, const float* Eh) {
    if (c->R == 4) ref_step_r<4>(c, Eh);
    else ref_step_r<6>(c, Eh);
}

int zero_state(erm2d_ctx* c, bool B) {
    const F2 f = state(c, B);
    const size_t n = c->ncell * nshot_of(c);
    for (float* a : {f.vx, f.vz, f.xx, f.zz, f.xz, f.tp, f.vxp, f.vzp}) if (fillf(c, a, n, 0.f)) return ERM_E_DEVICE;
    if (!B && !c->nb && c->psi && fillf(c, c->psi, 10 * c->ncell, 0.f)) return ERM_E_DEVICE;
    if (B && !c->nb && c->psiB && fillf(c, c->psiB, 10 * c->ncell, 0.f)) return ERM_E_DEVICE;
    for (int l = 0; l < c->L; ++l)
        for (int q = 0; q < 4; ++q) if (fillf(c, f.r[l][q], n, 0.f)) return ERM_E_DEVICE;
    return 0;
}
int zero_excitation(erm2d_ctx* c) {
    hipLaunchKernelGGL(k2_filli, dim3(512), dim3(256), 0, c->stream, c->ex_time, (long)c->ncell, -1);
    hipLaunchKernelGGL(k2_filli, dim3(512), dim3(256), 0, c->stream, c->ex_tp_time, (long)c->ncell, -1);
    if (fillf(c, c->ex_tp_amp, c->ncell, 0.f)) return ERM_E_DEVICE;
    for (auto* e : c->ex) if (fillf(c, e, c->ncell, 0.f)) return ERM_E_DEVICE;
    return 0;
}

float wavelet_at(const erm2d_ctx* c, int it) { return (it >= 0 && it < (int)c->wavelet.size()) ? c->wavelet[it] : 0.f; }

int run(erm2d_ctx* c, int pass, int it_begin, int nsteps, float* recbuf) {
    const int which = pass == ERM2D_PASS_MODEL ? 0 : 1;
    if (!c->mod[which].have) return c->fail(ERM_E_STATE, "%s model not loaded", which ? "background" : "true");
    c->last_model = which;
    const long comp = (long)c->p.lt * c->p.nrec;
    const int rg = (c->p.nrec + 255) / 256;
    c->qsign = pass == ERM2D_PASS_RECEIVER && c->p.qcomp ? -1.f : 1.f;
    for (int s = 0; s < nsteps; ++s) {
        if (pass == ERM2D_PASS_RECEIVER) {
            const int it = it_begin - s;
            if (it < 0 || it >= c->p.lt) return c->fail(ERM_E_PARAM, "time index out of range");
            if (c->p.nrec > 0)
                hipLaunchKernelGGL(k2_inject, dim3(rg), dim3(256), 0, c->stream, c->fA[0], c->fA[1], c->rec_idx,
                                   c->p.nrec, recbuf, comp, (long)it * c->p.nrec, 0L, 0L);
            do_stress(c, false, which, -1, 0.f, nullptr);
            do_velocity(c, false, which, 2, it, nullptr);
        } else {
            const int it = it_begin + s;
            do_stress(c, false, which, c->src, wavelet_at(c, it), nullptr);
            do_velocity(c, false, which, pass, it, nullptr);
            if (pass == ERM2D_PASS_MODEL && it < c->p.lt && c->p.nrec > 0)
                hipLaunchKernelGGL(k2_record, dim3(rg), dim3(256), 0, c->stream, c->fA[0], c->fA[1], c->rec_idx,
                                   c->p.nrec, recbuf, comp, (long)it * c->p.nrec, 0L, 0L);
        }
    }
    c->qsign = 1.f;
    HC(c, hipGetLastError());
    HC(c, hipStreamSynchronize(c->stream));
    return 0;
}

int ensure_lsrtm(erm2d_ctx* c) {
    if (!c->hist && dmal(c, &c->hist, (size_t)hist_steps(c) * 5 * c->mn)) return ERM_E_NOMEM;
    if (ls_strips(c)) {
        int rc;
        if ((rc = ensure_ring(c))) return rc;
        if (!c->ls_strip && dmal(c, &c->ls_strip, (size_t)c->p.lt * 8 * c->xc_nr)) return ERM_E_NOMEM;
    }
    if (!c->dm && dmal(c, &c->dm, 3 * c->mn)) return ERM_E_NOMEM;
    if (!c->grad && dmal(c, &c->grad, 3 * c->mn)) return ERM_E_NOMEM;
    return 0;
}

void free2(erm2d_ctx* c) {
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    for (int q = 0; q < 8; ++q) { F(c->fA[q]); F(c->fB[q]); }
    for (int l = 0; l < 3; ++l) for (int q = 0; q < 4; ++q) { F(c->rA[l][q]); F(c->rB[l][q]); }
    for (auto& w : c->W) F(w);
    F(c->dv); F(c->cu);
    for (auto& m : c->mod) { F(m.M); F(m.U); F(m.B); for (int l = 0; l < 3; ++l) { F(m.kP[l]); F(m.kS[l]); F(m.decay[l]); } }
    F(c->att); F(c->rec_idx); F(c->rec); F(c->rec2); F(c->ex_time); F(c->ex_tp_time); F(c->ex_tp_amp);
    F(c->psi); F(c->pcx); F(c->pcz); F(c->psiB);
    F(c->xc_img); F(c->xc_vote); F(c->xc_sign); F(c->xc_hist); F(c->xc_ring); F(c->xc_strip);
    for (auto& m : c->mod) { F(m.vp); F(m.vs); }
    for (auto& e : c->ex) F(e);
    for (auto& e : c->img) F(e);
    F(c->hist); F(c->ls_strip); F(c->dm); F(c->grad); F(c->illum);
    for (int q = 0; q < 8; ++q) { F(c->bt.A[q]); F(c->bt.B[q]); }
    for (int l = 0; l < 3; ++l) for (int q = 0; q < 4; ++q) { F(c->bt.rA[l][q]); F(c->bt.rB[l][q]); }
    for (auto& w : c->bt.W) F(w);
    F(c->bt.hist); F(c->bt.grad); F(c->bt.rec); F(c->bt.srcs); F(c->bt.strip);
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

int upload(erm2d_ctx* c, float* d, const std::vector<float>& h) {
    HC(c, hipMemcpyAsync(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HC(c, hipStreamSynchronize(c->stream));
    return 0;
}
// dense [nnx][nnz] host <-> pitched device
int up2(erm2d_ctx* c, float* d, const float* h) {
    HC(c, hipMemcpy2DAsync(d, (size_t)c->pz * 4, h, (size_t)c->nnz * 4, (size_t)c->nnz * 4, c->nnx,
                           hipMemcpyHostToDevice, c->stream));
    HC(c, hipStreamSynchronize(c->stream));
    return 0;
}
int down2(erm2d_ctx* c, float* h, const float* d) {
    HC(c, hipMemcpy2DAsync(h, (size_t)c->nnz * 4, d, (size_t)c->pz * 4, (size_t)c->nnz * 4, c->nnx,
                           hipMemcpyDeviceToHost, c->stream));
    HC(c, hipStreamSynchronize(c->stream));
    return 0;
}

}  // namespace

extern "C" {

const char* erm2d_last_error(const erm2d_ctx* c) { return c ? c->err.c_str() : g_err2.c_str(); }
size_t erm2d_params_size(void) { return sizeof(erm2d_params); }

int erm2d_params_default(erm2d_params* p) {
    if (!p) return ERM_E_PARAM;
    *p = erm2d_params{};
    p->fd_order = 8;
    p->lt_forward = -1;
    p->rec_dx = 1;
    return 0;
}

int erm2d_create(const erm2d_params* pin, int device, erm2d_ctx** out) {
    if (!pin || !out) { g_err2 = "null argument"; return ERM_E_PARAM; }
    *out = nullptr;
    const erm2d_params& p = *pin;
    int R = p.fd_order == 8 ? 4 : p.fd_order == 12 ? 6 : -1;
    if (R < 0) { g_err2 = "fd_order must be 8 or 12"; return ERM_E_PARAM; }
    if (p.nsls != 0 && p.nsls != 1 && p.nsls != 3) { g_err2 = "nsls must be 0, 1 or 3"; return ERM_E_PARAM; }
    if (p.nx <= 0 || p.nz <= 0 || p.lt <= 0 || p.dt <= 0 || p.dx <= 0 || p.dz <= 0) { g_err2 = "bad grid"; return ERM_E_PARAM; }
    if (p.bu < 0 || p.bd < 0 || p.bl < 0 || p.br < 0) { g_err2 = "negative boundary width"; return ERM_E_PARAM; }
    if (p.boundary != ERM_BOUNDARY_DAMP && p.boundary != ERM_BOUNDARY_CPML) { g_err2 = "boundary must be 0 (damping) or 1 (CPML)"; return ERM_E_PARAM; }
    if (p.inversion_para < 0 || p.inversion_para > 3) { g_err2 = "inversion_para must be 1, 2 or 3 (0: 2)"; return ERM_E_PARAM; }
    if (p.adjoint_mode != 0 && p.adjoint_mode != 1) { g_err2 = "adjoint_mode must be 0 (exact transpose) or 1 (reference)"; return ERM_E_PARAM; }
    if (p.tile_rows != 0 && p.tile_rows != TXR && p.tile_rows != TXR_TALL) { g_err2 = "tile_rows must be 0, 4 or 8"; return ERM_E_PARAM; }
    if (p.adjoint_source != ERM2D_SRC_HISTORY && p.adjoint_source != ERM2D_SRC_STRIPS) {
        g_err2 = "adjoint_source must be 0 (history) or 2 (strips)";
        return ERM_E_PARAM;
    }
    if (p.adjoint_source == ERM2D_SRC_STRIPS && (p.nsls != 0 || p.adjoint_mode != 0)) {
        g_err2 = "adjoint_source 2 (strips) needs an elastic medium (nsls = 0) and adjoint_mode 0";
        return ERM_E_PARAM;
    }
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) { g_err2 = "no HIP device"; return ERM_E_DEVICE; }
    if (device < 0 || device >= nd) { g_err2 = "device out of range"; return ERM_E_PARAM; }
    erm2d_ctx* c = new erm2d_ctx();
    c->p = p;
    if (c->p.lt_forward < 0) c->p.lt_forward = p.lt;
    c->device = device;
    c->R = R;
    c->L = p.nsls;
    c->nnx = p.nx + p.bl + p.br;
    c->nnz = p.nz + p.bu + p.bd;
    if (c->nnx < 2 * R + 1 || c->nnz < 2 * R + 1) { delete c; g_err2 = "padded grid smaller than the stencil"; return ERM_E_SIZE; }
    c->pz = (c->nnz + 31) / 32 * 32;
    c->ncell = (size_t)c->nnx * c->pz;
    c->mn = (size_t)p.nx * p.nz;
    for (int m = 0; m <= R; ++m) c->coef.c[m] = (float)(R == 6 ? kC6[m] : kC4[m]);
    c->dtr = (float)(p.dt / 1000.0);
    c->cx = (float)(p.dt / 1000.0 / p.dx);
    c->cz = (float)(p.dt / 1000.0 / p.dz);
#define CK2(x) do { int r_ = (x); if (r_) { g_err2 = c->err; free2(c); delete c; return r_; } } while (0)
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_err2 = "stream create failed"; delete c; return ERM_E_DEVICE;
    }
    for (auto& e : c->ev) if (hipEventCreate(&e) != hipSuccess) { g_err2 = "event create failed"; free2(c); delete c; return ERM_E_DEVICE; }
    if (resolve_scratch(device)) { g_err2 = "scratch symbol address lookup failed"; free2(c); delete c; return ERM_E_DEVICE; }
    for (int q = 0; q < 8; ++q) { CK2(dmal(c, &c->fA[q], c->ncell)); CK2(dmal(c, &c->fB[q], c->ncell)); }
    for (int l = 0; l < c->L; ++l)
        for (int q = 0; q < 4; ++q) { CK2(dmal(c, &c->rA[l][q], c->ncell)); CK2(dmal(c, &c->rB[l][q], c->ncell)); }
    for (auto& w : c->W) CK2(dmal(c, &w, c->ncell));
    CK2(dmal(c, &c->dv, c->ncell));
    CK2(dmal(c, &c->cu, c->ncell));
    CK2(dmal(c, &c->att, c->ncell));
    c->cpml = p.boundary == ERM_BOUNDARY_CPML;
    if (c->cpml) {
        CK2(dmal(c, &c->psi, 10 * c->ncell));
        CK2(dmal(c, &c->pcx, 4 * (size_t)c->nnx));
        CK2(dmal(c, &c->pcz, 4 * (size_t)c->nnz));
        CK2(fillf(c, c->pcx, 4 * (size_t)c->nnx, 0.f));
        CK2(fillf(c, c->pcz, 4 * (size_t)c->nnz, 0.f));
    }
    CK2(dmal(c, &c->ex_time, c->ncell));
    CK2(dmal(c, &c->ex_tp_time, c->ncell));
    CK2(dmal(c, &c->ex_tp_amp, c->ncell));
    for (auto& e : c->ex) CK2(dmal(c, &e, c->ncell));
    for (auto& e : c->img) CK2(dmal(c, &e, c->ncell));
    // damping: make_attenuation_new (zzzzz:238-325), coe * s^3
    {
        std::vector<double> ax(c->nnx, 0), az(c->nnz, 0);
        auto face = [&](int b, bool rising, std::vector<double>& axis, int start) {
            for (int t = 0; t < b; ++t) {
                double s = rising ? t / (b - 1.0) : (b - 1.0 - t) / (b - 1.0);
                axis[start + t] += p.coe_attenuation * std::pow(s, 3.0);
            }
        };
        if (!c->cpml) {  // the C-PML replaces the damping
            if (p.bu > 1) face(p.bu, false, az, 0);
            if (p.bd > 1) face(p.bd, true, az, p.bu + p.nz);
            if (p.bl > 1) face(p.bl, false, ax, 0);
            if (p.br > 1) face(p.br, true, ax, p.bl + p.nx);
        }
        std::vector<float> h(c->ncell, 0.f);
        for (int i = 0; i < c->nnx; ++i)
            for (int k = 0; k < c->nnz; ++k) h[(size_t)i * c->pz + k] = (float)(ax[i] + az[k]);
        CK2(upload(c, c->att, h));
    }
    // make_ricker_new (zzzzz:136-157)
    {
        const double PI = 3.141592653589793;
        const float dts = (float)(p.dt / 1000.0);
        float B = (float)(std::sqrt(6.0) / (PI * p.freq));
        int N = (int)std::ceil(1.35 * (double)B / (double)dts);
        c->wavelet.assign(2 * N + 1, 0.f);
        c->wavelet[N] = 1.f;
        for (int i = 1; i <= N; ++i) {
            float t = dts * (float)i, u = (float)(2.0 * std::sqrt(6.0)) * t / B;
            c->wavelet[N + i] = c->wavelet[N - i] = 0.5f * (2.0f - u * u) * std::exp(-u * u / 4.0f);
        }
    }
    if (p.nrec > 0) {
        std::vector<long> idx(p.nrec);
        for (int q = 0; q < p.nrec; ++q) {
            int i = p.bl + p.rec_x0 + q * p.rec_dx, k = p.bu + p.rec_z;
            if (i < 0 || i >= c->nnx || k < 0 || k >= c->nnz) { g_err2 = "receiver outside the grid"; free2(c); delete c; return ERM_E_PARAM; }
            idx[q] = (long)i * c->pz + k;
        }
        CK2(dmal(c, &c->rec_idx, idx.size()));
        if (hipMemcpy(c->rec_idx, idx.data(), idx.size() * sizeof(long), hipMemcpyHostToDevice) != hipSuccess) {
            g_err2 = "receiver upload failed"; free2(c); delete c; return ERM_E_DEVICE;
        }
        CK2(dmal(c, &c->rec, (size_t)2 * p.lt * p.nrec));
        CK2(dmal(c, &c->rec2, (size_t)2 * p.lt * p.nrec));
    }
    CK2(zero_state(c, false));
    CK2(zero_state(c, true));
    for (auto& w : c->W) CK2(fillf(c, w, c->ncell, 0.f));
    CK2(zero_excitation(c));
    for (auto& e : c->img) CK2(fillf(c, e, c->ncell, 0.f));
    CK2(erm2d_set_shot(c, p.shot_x, p.shot_z));
    if (hipStreamSynchronize(c->stream) != hipSuccess) { g_err2 = "init failed"; free2(c); delete c; return ERM_E_DEVICE; }
#undef CK2
    *out = c;
    return 0;
}

int erm2d_destroy(erm2d_ctx* c) {
    if (!c) return 0;
    for (auto* w : c->workers) erm2d_destroy(w);
    c->workers.clear();
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free2(c);
    delete c;
    return 0;
}

int erm2d_get_dims(const erm2d_ctx* c, int* nnx, int* nnz) {
    if (!c) return ERM_E_PARAM;
    if (nnx) *nnx = c->nnx;
    if (nnz) *nnz = c->nnz;
    return 0;
}

int erm2d_load_model(erm2d_ctx* c, int which, const float* vp, const float* vs, const float* rho, const float* qp,
                     const float* qs) {
    if (!c || which < 0 || which > 1 || !vp || !vs || !rho) return c ? c->fail(ERM_E_PARAM, "bad arguments") : ERM_E_PARAM;
    if (c->L > 0 && (!qp || !qs)) return c->fail(ERM_E_PARAM, "qp/qs required when nsls > 0");
    HC(c, hipSetDevice(c->device));
    const erm2d_params& p = c->p;
    std::vector<float> hM(c->ncell, 0.f), hU(c->ncell, 0.f), hB(c->ncell, 0.f), hvp(c->ncell, 0.f), hvs(c->ncell, 0.f);
    std::vector<float> hkP[3], hkS[3], hd[3];
    for (int l = 0; l < c->L; ++l) { hkP[l].assign(c->ncell, 0.f); hkS[l].assign(c->ncell, 0.f); hd[l].assign(c->ncell, 0.f); }
    for (int i = 0; i < c->nnx; ++i)
        for (int k = 0; k < c->nnz; ++k) {
            int ix = std::min(std::max(i - p.bl, 0), p.nx - 1), iz = std::min(std::max(k - p.bu, 0), p.nz - 1);
            size_t s = (size_t)ix * p.nz + iz, n = (size_t)i * c->pz + k;
            double a = vp[s], b = vs[s], r = rho[s];
            if (!(a > 0) || b < 0 || !(r > 0)) return c->fail(ERM_E_PARAM, "non-physical model value at (%d,%d)", ix, iz);
            Sls q = sls_params(c->L, p.freq, c->L ? qp[s] : 1.0, c->L ? qs[s] : 1.0);
            if (c->L && (!(qp[s] > 0) || !(qs[s] > 0))) return c->fail(ERM_E_PARAM, "Q must be positive");
            hM[n] = (float)(r * a * a * q.fM);
            hU[n] = (float)(r * b * b * q.fU);
            hB[n] = (float)(1.0 / r);
            hvp[n] = vp[s];
            hvs[n] = vs[s];
            for (int l = 0; l < c->L; ++l) {
                hkP[l][n] = (float)(r * a * a * q.gP[l]);
                hkS[l][n] = (float)(r * b * b * q.gS[l]);
                hd[l][n] = (float)(p.dt / 1000.0 / q.taus[l]);
            }
        }
    auto& m = c->mod[which];
    if (!m.M) {
        if (dmal(c, &m.M, c->ncell) || dmal(c, &m.U, c->ncell) || dmal(c, &m.B, c->ncell) ||
            dmal(c, &m.vp, c->ncell) || dmal(c, &m.vs, c->ncell))
            return ERM_E_NOMEM;
        for (int l = 0; l < c->L; ++l)
            if (dmal(c, &m.kP[l], c->ncell) || dmal(c, &m.kS[l], c->ncell) || dmal(c, &m.decay[l], c->ncell)) return ERM_E_NOMEM;
    }
    if (upload(c, m.M, hM) || upload(c, m.U, hU) || upload(c, m.B, hB) || upload(c, m.vp, hvp) || upload(c, m.vs, hvs))
        return ERM_E_DEVICE;
    for (int l = 0; l < c->L; ++l)
        if (upload(c, m.kP[l], hkP[l]) || upload(c, m.kS[l], hkS[l]) || upload(c, m.decay[l], hd[l])) return ERM_E_DEVICE;
    m.have = true;
    for (size_t n = 0; n < c->mn; ++n) c->vmax_loaded = std::max(c->vmax_loaded, (double)vp[n]);
    if (c->cpml) {
        // C-PML coefficients (Roden & Gedney 2000; Komatitsch & Martin 2007), kappa = 1, as in 3D:
        // d = d0 r^2, d0 = -3 vmax ln(Rc) / (2 L h), alpha = pi f0 (1 - r), b = exp(-(d+alpha) dt),
        // a = d (b-1)/(d+alpha); r = depth into the layer / L, L = width - R; full and half
        // (x - 1/2, z + 1/2) positions
        const double vmax = p.cpml_vmax > 0 ? p.cpml_vmax : c->vmax_loaded;
        const double Rc = p.cpml_r > 0 ? p.cpml_r : 1e-3, dts = p.dt / 1000.0;
        const int n2[2] = {c->nnx, c->nnz}, lo[2] = {p.bl, p.bu}, hi[2] = {p.br, p.bd};
        const double h[2] = {p.dx, p.dz}, half_shift[2] = {-0.5, +0.5};
        for (int a = 0; a < 2; ++a) {
            std::vector<float> pc(4 * (size_t)n2[a], 0.f);
            for (int f = 0; f < 2; ++f)
                for (int i = 0; i < n2[a]; ++i) {
                    double x = i + (f ? half_shift[a] : 0.0), sd = 0, Lw = 0;
                    pc[(2 * f + 1) * (size_t)n2[a] + i] = 1.f;
                    if (lo[a] > 0 && x < lo[a]) { sd = lo[a] - x; Lw = std::max(lo[a] - c->R, 1); }
                    else if (hi[a] > 0 && x > n2[a] - hi[a] - 1) { sd = x - (n2[a] - hi[a] - 1); Lw = std::max(hi[a] - c->R, 1); }
                    if (Lw <= 0) continue;
                    const double r = sd / Lw, d0 = -3.0 * vmax * std::log(Rc) / (2.0 * Lw * h[a]);
                    const double d = d0 * r * r, al = M_PI * p.freq * (1.0 - std::min(r, 1.0));
                    const double bb = std::exp(-(d + al) * dts);
                    pc[(2 * f + 1) * (size_t)n2[a] + i] = (float)bb;
                    pc[(2 * f) * (size_t)n2[a] + i] = (float)(d > 0 ? d * (bb - 1.0) / (d + al) : 0.0);
                }
            if (upload(c, a == 0 ? c->pcx : c->pcz, pc)) return ERM_E_DEVICE;
        }
    }
    if (which == 1) {
        const float* src[5] = {vp, vs, rho, c->L ? qp : nullptr, c->L ? qs : nullptr};
        for (int a = 0; a < 5; ++a) {
            if (src[a]) c->bg_host[a].assign(src[a], src[a] + c->mn);
            else c->bg_host[a].clear();
        }
        for (auto* w : c->workers) {
            int rc = erm2d_load_model(w, 1, vp, vs, rho, qp, qs);
            if (rc) return c->fail(rc, "worker: %s", w->err.c_str());
        }
    }
    return 0;
}

int erm2d_set_workers(erm2d_ctx* c, int n) {
    if (!c || n < 1 || n > 16) return c ? c->fail(ERM_E_PARAM, "workers must be 1..16") : ERM_E_PARAM;
    while ((int)c->workers.size() > n - 1) {
        erm2d_destroy(c->workers.back());
        c->workers.pop_back();
    }
    while ((int)c->workers.size() < n - 1) {
        erm2d_ctx* w = nullptr;
        int rc = erm2d_create(&c->p, c->device, &w);
        if (rc) return c->fail(rc, "worker create: %s", g_err2.c_str());
        c->workers.push_back(w);
        if (c->mod[1].have) {
            auto P = [&](int a) { return c->bg_host[a].empty() ? nullptr : c->bg_host[a].data(); };
            rc = erm2d_load_model(w, 1, P(0), P(1), P(2), P(3), P(4));
            if (rc) return c->fail(rc, "worker load: %s", w->err.c_str());
        }
    }
    return 0;
}

int erm2d_set_cg_monitor(erm2d_ctx* c, erm2d_cg_monitor_fn fn, void* user) {
    if (!c) return ERM_E_PARAM;
    c->monitor = fn;
    c->monitor_user = user;
    return 0;
}

int erm2d_set_shot(erm2d_ctx* c, int sx, int sz) {
    if (!c) return ERM_E_PARAM;
    int i = c->p.bl + sx, k = c->p.bu + sz;
    if (i < c->R || i >= c->nnx - c->R || k < c->R || k >= c->nnz - c->R)
        return c->fail(ERM_E_PARAM, "source (%d,%d) is not an updated cell", i, k);
    c->p.shot_x = sx;
    c->p.shot_z = sz;
    c->src = (long)i * c->pz + k;
    return 0;
}

int erm2d_zero(erm2d_ctx* c) {
    if (!c) return ERM_E_PARAM;
    if (zero_state(c, false)) return ERM_E_DEVICE;
    HC(c, hipStreamSynchronize(c->stream));
    return 0;
}

int erm2d_step(erm2d_ctx* c, int pass, int it_begin, int nsteps) {
    if (!c) return ERM_E_PARAM;
    HC(c, hipSetDevice(c->device));
    if (pass < 0 || pass > 2) return c->fail(ERM_E_PARAM, "bad pass");
    return run(c, pass, it_begin, nsteps, c->rec);
}

int erm2d_get_field(erm2d_ctx* c, int field, float* dst) {
    if (!c || !dst) return ERM_E_PARAM;
    HC(c, hipSetDevice(c->device));
    if (field >= 0 && field < 8) return down2(c, dst, c->fA[field]);
    if (field == ERM2D_DIV || field == ERM2D_CURL || field == ERM2D_PDIV || field == ERM2D_SCURL) {
        G2 g = geo(c);
        const bool scaled = field == ERM2D_PDIV || field == ERM2D_SCURL;
        const auto& md = c->mod[c->last_model];
        if (scaled && !md.have) return c->fail(ERM_E_STATE, "model of the last pass not loaded");
        const float* svp = scaled ? md.vp : nullptr;
        const float* svs = scaled ? md.vs : nullptr;
        if (c->R == 4)
            hipLaunchKernelGGL((k2_decom<4, TXR>), grid2(c), dim3(TZ, TXR), 0, c->stream, c->fA[0], c->fA[1], c->dv,
                               c->cu, g, c->coef, (float)(1.0 / c->p.dx), (float)(1.0 / c->p.dz), svp, svs);
        else
            hipLaunchKernelGGL((k2_decom<6, TXR>), grid2(c), dim3(TZ, TXR), 0, c->stream, c->fA[0], c->fA[1], c->dv,
                               c->cu, g, c->coef, (float)(1.0 / c->p.dx), (float)(1.0 / c->p.dz), svp, svs);
        HC(c, hipGetLastError());
        return down2(c, dst, field == ERM2D_DIV || field == ERM2D_PDIV ? c->dv : c->cu);
    }
    if (field == ERM2D_EX_TIME || field == ERM2D_EX_TP_TIME) {
        std::vector<int> t(c->ncell);
        HC(c, hipMemcpyAsync(t.data(), field == ERM2D_EX_TIME ? c->ex_time : c->ex_tp_time, c->ncell * sizeof(int),
                             hipMemcpyDeviceToHost, c->stream));
        HC(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < c->nnx; ++i)
            for (int k = 0; k < c->nnz; ++k) dst[(size_t)i * c->nnz + k] = (float)t[(size_t)i * c->pz + k];
        return 0;
    }
    if (field == ERM2D_EX_AMP) return down2(c, dst, c->ex[0]);
    if (field == ERM2D_EX_TP_AMP) return down2(c, dst, c->ex_tp_amp);
    if (field >= ERM2D_EX_TP && field <= ERM2D_R_VZP) return down2(c, dst, c->ex[1 + field - ERM2D_EX_TP]);
    if (field >= ERM2D_IMG_TP && field <= ERM2D_IMG_PS) return down2(c, dst, c->img[field - ERM2D_IMG_TP]);
    return c->fail(ERM_E_PARAM, "unknown field %d", field);
}

int erm2d_set_field(erm2d_ctx* c, int field, const float* src) {
    if (!c || !src) return ERM_E_PARAM;
    if (field < 0 || field >= 8) return c->fail(ERM_E_PARAM, "field %d is not writable", field);
    return up2(c, c->fA[field], src);
}

int erm2d_forward(erm2d_ctx* c, float* rec) {
    if (!c) return ERM_E_PARAM;
    HC(c, hipSetDevice(c->device));
    int rc;
    if ((rc = zero_state(c, false))) return rc;
    if ((rc = run(c, ERM2D_PASS_MODEL, 0, c->p.lt, c->rec))) return rc;
    if (rec && c->p.nrec > 0) {
        HC(c, hipMemcpyAsync(rec, c->rec, (size_t)2 * c->p.lt * c->p.nrec * 4, hipMemcpyDeviceToHost, c->stream));
        HC(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

static int cut_model(erm2d_ctx* c, const float* d, float* out) {
    std::vector<float> full((size_t)c->nnx * c->nnz);
    int rc = down2(c, full.data(), d);
    if (rc) return rc;
    for (int ix = 0; ix < c->p.nx; ++ix)
        for (int iz = 0; iz < c->p.nz; ++iz)
            out[(size_t)ix * c->p.nz + iz] = full[(size_t)(ix + c->p.bl) * c->nnz + iz + c->p.bu];
    return 0;
}

}  // extern "C"

namespace {
int timed_ms(erm2d_ctx* c);
void decom_into(erm2d_ctx* c, bool B, float* dv, float* cu) {
    const F2 f = state(c, B);
    const G2 g = geo(c);
    if (c->R == 4)
        hipLaunchKernelGGL((k2_decom<4, TXR>), grid2(c), dim3(TZ, TXR), 0, c->stream, f.vx, f.vz, dv, cu, g, c->coef,
                           (float)(1.0 / c->p.dx), (float)(1.0 / c->p.dz), (const float*)nullptr, (const float*)nullptr);
    else
        hipLaunchKernelGGL((k2_decom<6, TXR>), grid2(c), dim3(TZ, TXR), 0, c->stream, f.vx, f.vz, dv, cu, g, c->coef,
                           (float)(1.0 / c->p.dx), (float)(1.0 / c->p.dz), (const float*)nullptr, (const float*)nullptr);
}
// ring of the strips strategy: cells outside the model region shrunk by one
// cell within R + 1 of it (R for the reversed stencils, one more for the
// div / curl of the model region's border cells)
std::vector<long> xcorr_ring(const erm2d_ctx* c) {
    const int i0 = c->p.bl + 1, i1 = c->p.bl + c->p.nx - 1, k0 = c->p.bu + 1, k1 = c->p.bu + c->p.nz - 1;
    std::vector<long> out;
    for (int i = std::max(i0 - c->R - 1, 0); i < std::min(i1 + c->R + 1, c->nnx); ++i)
        for (int k = std::max(k0 - c->R - 1, 0); k < std::min(k1 + c->R + 1, c->nnz); ++k)
            if (i < i0 || i >= i1 || k < k0 || k >= k1) out.push_back((long)i * c->pz + k);
    return out;
}
int ensure_ring(erm2d_ctx* c) {
    if (c->xc_ring) return 0;
    std::vector<long> rg = xcorr_ring(c);
    c->xc_nr = (long)rg.size();
    if (dmal(c, &c->xc_ring, std::max<size_t>(rg.size(), 1))) return ERM_E_NOMEM;
    if (!rg.empty())
        HC(c, hipMemcpyAsync(c->xc_ring, rg.data(), rg.size() * sizeof(long), hipMemcpyHostToDevice, c->stream));
    HC(c, hipStreamSynchronize(c->stream));
    return 0;
}
// save or restore the ring of state A's stress side (sxx, szz, sxz, tp) or
// velocity side (vx, vz, vxp, vzp) for every shot of the running operator;
// buf is shot 0's [4][nr] block, shot z's sits bst floats further (restore
// with a null buf writes zeros)
void ring_io(erm2d_ctx* c, bool vside, float* buf, int save, long bst = 0) {
    if (c->xc_nr == 0) return;  // the model region fills the grid: no ring
    const F2 f = state(c, false);
    float* a[4] = {vside ? f.vx : f.xx, vside ? f.vz : f.zz, vside ? f.vxp : f.xz, vside ? f.vzp : f.tp};
    const dim3 grid((unsigned)std::min<long>((c->xc_nr + 255) / 256, 1024), nshot_of(c));
    hipLaunchKernelGGL(k2x_ring, grid, dim3(256), 0, c->stream, a[0], a[1], a[2], a[3], c->xc_ring, c->xc_nr, buf, save,
                       c->nb ? (long)c->ncell : 0L, buf ? bst : 0L);
}
}  // namespace

extern "C" {

int erm2d_rtm_xcorr(erm2d_ctx* c, const float* rec, int strategy, float* images) {
    if (!c) return ERM_E_PARAM;
    if (strategy < 0 || strategy > 2) return c->fail(ERM_E_PARAM, "strategy must be 0 (history), 1 (reverse) or 2 (strips)");
    if (strategy == 2 && c->L != 0) return c->fail(ERM_E_PARAM, "the strips strategy needs an elastic medium (nsls = 0)");
    HC(c, hipSetDevice(c->device));
    if (!c->mod[1].have) return c->fail(ERM_E_STATE, "background model not loaded");
    const int ltf = std::min(c->p.lt_forward < 0 ? c->p.lt : c->p.lt_forward, c->p.lt);
    const size_t mn = c->mn;
    if (!c->xc_img && dmal(c, &c->xc_img, 5 * mn)) return ERM_E_NOMEM;
    if (!c->xc_vote && dmal(c, &c->xc_vote, c->ncell)) return ERM_E_NOMEM;
    if (!c->xc_sign && dmal(c, &c->xc_sign, c->ncell)) return ERM_E_NOMEM;
    if (c->cpml && !c->psiB && dmal(c, &c->psiB, 10 * c->ncell)) return ERM_E_NOMEM;
    if (strategy == 0 && c->xc_hist_n < (size_t)ltf * 3 * mn) {
        if (c->xc_hist) { (void)hipFree(c->xc_hist); c->xc_hist = nullptr; c->xc_hist_n = 0; }
        if (dmal(c, &c->xc_hist, (size_t)ltf * 3 * mn)) return ERM_E_NOMEM;
        c->xc_hist_n = (size_t)ltf * 3 * mn;
    }
    if (strategy == 2) {
        if (const int e = ensure_ring(c)) return e;
        const size_t need = (size_t)ltf * 8 * c->xc_nr;
        if (c->xc_strip_n < need) {
            if (c->xc_strip) { (void)hipFree(c->xc_strip); c->xc_strip = nullptr; c->xc_strip_n = 0; }
            if (dmal(c, &c->xc_strip, need)) return ERM_E_NOMEM;
            c->xc_strip_n = need;
        }
    }
    if (rec && c->p.nrec > 0)
        HC(c, hipMemcpyAsync(c->rec, rec, (size_t)2 * c->p.lt * c->p.nrec * 4, hipMemcpyHostToDevice, c->stream));
    HC(c, hipEventRecord(c->ev[0], c->stream));
    int rc;
    if ((rc = zero_state(c, false)) || (rc = zero_state(c, true))) return rc;
    if ((rc = fillf(c, c->xc_img, 5 * mn, 0.f))) return rc;
    hipLaunchKernelGGL(k2_filli, dim3(512), dim3(256), 0, c->stream, c->xc_sign, (long)c->ncell, 0);
    c->last_model = 1;
    const int which = 1;
    const long nr8 = 8 * c->xc_nr;
    float* sdv = c->W[0];   // source div (padded grid)
    float* prod = c->W[1];  // div_s curl_r
    // source pass (state A)
    for (int it = 0; it < ltf; ++it) {
        do_stress(c, false, which, c->src, wavelet_at(c, it), nullptr);
        if (strategy == 2) ring_io(c, false, c->xc_strip + it * nr8, 1);
        do_velocity(c, false, which, 0, it, nullptr);
        if (strategy == 2) ring_io(c, true, c->xc_strip + it * nr8 + 4 * c->xc_nr, 1);
        if (strategy == 0) {
            decom_into(c, false, sdv, prod);
            hipLaunchKernelGGL(k2x_save_hist, dim3((unsigned)std::min<size_t>((mn + 255) / 256, 2048)), dim3(256), 0,
                               c->stream, geo(c), c->fA[6], c->fA[7], sdv, c->xc_hist + (size_t)it * 3 * mn);
        }
    }
    // receiver pass (state B), imaging, source reversal
    const long comp = (long)c->p.lt * c->p.nrec;
    const int rg = (c->p.nrec + 255) / 256;
    const F2 fb = state(c, true);
    const dim3 gpt = grid_pt(c), gpol((c->nnz + TZ - 1) / TZ, (c->nnx + 3) / 4);
    for (int it = c->p.lt - 1; it >= 0; --it) {
        if (c->p.nrec > 0)
            hipLaunchKernelGGL(k2_inject, dim3(rg), dim3(256), 0, c->stream, fb.vx, fb.vz, c->rec_idx, c->p.nrec, c->rec,
                               comp, (long)it * c->p.nrec, 0L, 0L);
        c->qsign = c->p.qcomp ? -1.f : 1.f;
        do_stress(c, true, which, -1, 0.f, nullptr);
        c->qsign = 1.f;
        do_velocity(c, true, which, 0, it, nullptr);
        if (it >= ltf) continue;
        decom_into(c, true, c->dv, c->cu);
        const float *sx, *sz, *sd;
        int hist = 0;
        if (strategy == 0) {
            const float* h = c->xc_hist + (size_t)it * 3 * mn;
            sx = h; sz = h + mn; sd = h + 2 * mn; hist = 1;
        } else {
            decom_into(c, false, sdv, prod);
            sx = c->fA[6]; sz = c->fA[7]; sd = sdv;
        }
        hipLaunchKernelGGL(k2x_image, gpt, dim3(256), 0, c->stream, geo(c), sx, sz, sd, hist, fb, c->dv, c->cu,
                           c->xc_img, c->xc_vote, prod, c->R);
        hipLaunchKernelGGL(k2x_polarity, gpol, dim3(TZ, 4), 0, c->stream, geo(c), c->xc_vote, c->xc_sign, prod,
                           c->xc_img, c->R);
        if (strategy != 0) {  // undo source step it (rfwd_..._old)
            if (strategy == 2) ring_io(c, false, c->xc_strip + it * nr8, 0);
            c->dsgn = -1.f;
            do_velocity(c, false, which, 0, it, nullptr);
            c->dsgn = 1.f;
            if (strategy == 2) ring_io(c, true, it > 0 ? c->xc_strip + (it - 1) * nr8 + 4 * c->xc_nr : nullptr, 0);
            c->dsgn = -1.f;
            c->qsign = -1.f;
            do_stress(c, false, which, c->src, -wavelet_at(c, it), nullptr);
            c->dsgn = 1.f;
            c->qsign = 1.f;
        }
    }
    HC(c, hipGetLastError());
    if ((rc = timed_ms(c))) return rc;
    if (images) {
        HC(c, hipMemcpyAsync(images, c->xc_img, 5 * mn * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HC(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int erm2d_rtm(erm2d_ctx* c, const float* rec, float* tp, float* pp, float* ps) {
    if (!c) return ERM_E_PARAM;
    HC(c, hipSetDevice(c->device));
    if (rec && c->p.nrec > 0)
        HC(c, hipMemcpyAsync(c->rec, rec, (size_t)2 * c->p.lt * c->p.nrec * 4, hipMemcpyHostToDevice, c->stream));
    HC(c, hipEventRecord(c->ev[0], c->stream));
    int rc;
    if ((rc = zero_state(c, false)) || (rc = zero_excitation(c))) return rc;
    if ((rc = run(c, ERM2D_PASS_SOURCE, 0, c->p.lt_forward, c->rec))) return rc;
    if ((rc = zero_state(c, false))) return rc;
    for (auto& e : c->img) if (fillf(c, e, c->ncell, 0.f)) return ERM_E_DEVICE;
    if ((rc = run(c, ERM2D_PASS_RECEIVER, c->p.lt - 1, c->p.lt, c->rec))) return rc;
    HC(c, hipEventRecord(c->ev[1], c->stream));
    HC(c, hipEventSynchronize(c->ev[1]));
    float ms = 0;
    HC(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->last_ms = ms;
    float* outs[3] = {tp, pp, ps};
    for (int a = 0; a < 3; ++a)
        if (outs[a] && (rc = cut_model(c, c->img[a], outs[a]))) return rc;
    return 0;
}

}  // extern "C"

namespace {
// One Born application over the shots of the running operator (one, or nb
// batched along z): background and Born states stepped together, the Born
// sources from the background's strain increments and forces of the step
// (one step of E/F per shot in the history buffer); records to rec_of(c).
// c->dm holds the perturbation.
int born_core(erm2d_ctx* c) {
    int rc;
    if ((rc = zero_state(c, false)) || (rc = zero_state(c, true))) return rc;
    const long comp = (long)c->p.lt * c->p.nrec;
    const int rg = (c->p.nrec + 255) / 256;
    float* Eh = hist_of(c);
    float* Fh = Eh + 3 * c->mn;
    float* rec = rec_of(c);
    const long sst = c->nb ? (long)c->ncell : 0, rst = c->nb ? 2 * comp : 0;
    const F2 fb = state(c, true);
    for (int it = 0; it < c->p.lt; ++it) {
        do_stress(c, false, 1, c->src, wavelet_at(c, it), Eh);
        do_stress(c, true, 1, -1, 0.f, nullptr);
        born_stress_src(c, Eh);
        do_velocity(c, false, 1, 0, it, Fh);
        do_velocity(c, true, 1, 0, it, nullptr);
        hipLaunchKernelGGL(k2_born_velocity_src, grid_pt(c), dim3(256), 0, c->stream, fb, model(c, 1), geo(c), c->dm, Fh);
        if (c->p.nrec > 0)
            hipLaunchKernelGGL(k2_record, dim3(rg, nshot_of(c)), dim3(256), 0, c->stream, fb.vx, fb.vz, c->rec_idx,
                               c->p.nrec, rec, comp, (long)it * c->p.nrec, sst, rst);
    }
    HC(c, hipGetLastError());
    return 0;
}

// One adjoint application over the shots of the running operator: the
// background pass stores E and F of every step, then the transposed Born
// step runs backward from the residual records in rec_of(c); per-shot
// gradients in grad_of(c), summed into c->grad and mapped to the inverted
// parameters.  With adjoint_source strips the background pass stores only
// the ring strips of every half step, and before each transposed step the
// background steps back over that step (velocity, then stress, with the
// derivative terms and the source negated, rfwd_..._old), the ring restored
// from the strips: those reversed kernels recompute F and E of the step into
// the one-step history slot.
int adjoint_ref_core(erm2d_ctx* c, bool mapped);
int adjoint_core(erm2d_ctx* c) {
    if (c->p.adjoint_mode == 1) return adjoint_ref_core(c, true);
    int rc;
    const long comp = (long)c->p.lt * c->p.nrec;
    const int rg = (c->p.nrec + 255) / 256;
    const int ns = nshot_of(c);
    if ((rc = zero_state(c, false)) || (rc = zero_state(c, true))) return rc;
    if ((rc = fillf(c, grad_of(c), 3 * c->mn * ns, 0.f))) return rc;
    for (int q = 0; q < 3; ++q) if ((rc = fillf(c, w_of(c)[q], c->ncell * ns, 0.f))) return rc;
    float* hist = hist_of(c);
    float* rec = rec_of(c);
    const long sst = c->nb ? (long)c->ncell : 0, rst = c->nb ? 2 * comp : 0;
    const F2 fb = state(c, true);
    const size_t step = 5 * c->mn;
    const bool strips = ls_strips(c);
    float* strip = c->nb ? c->bt.strip : c->ls_strip;
    const long nr4 = 4 * c->xc_nr, bst = (long)c->p.lt * 2 * nr4;  // per-shot strip stride
    for (int it = 0; it < c->p.lt; ++it) {  // background, storing E and F per step, or the strips
        float* h = strips ? nullptr : hist + (size_t)it * step;
        do_stress(c, false, 1, c->src, wavelet_at(c, it), h);
        if (strips) ring_io(c, false, strip + 2 * it * nr4, 1, bst);
        do_velocity(c, false, 1, 0, it, h ? h + 3 * c->mn : nullptr);
        if (strips) ring_io(c, true, strip + (2 * it + 1) * nr4, 1, bst);
    }
    for (int it = c->p.lt - 1; it >= 0; --it) {
        const float* h = strips ? hist : hist + (size_t)it * step;
        if (strips) {  // background back over step it: F, then E of the step into the one-step slot
            ring_io(c, false, strip + 2 * it * nr4, 0, bst);
            c->dsgn = -1.f;
            do_velocity(c, false, 1, 0, it, hist + 3 * c->mn);
            ring_io(c, true, it > 0 ? strip + (2 * it - 1) * nr4 : nullptr, 0, bst);
            do_stress(c, false, 1, c->src, -wavelet_at(c, it), hist);
            c->dsgn = 1.f;
        }
        if (c->p.nrec > 0)
            hipLaunchKernelGGL(k2_inject, dim3(rg, ns), dim3(256), 0, c->stream, fb.vx, fb.vz, c->rec_idx, c->p.nrec,
                               rec, comp, (long)it * c->p.nrec, sst, rst);
        adj_v(c, h + 3 * c->mn);
        adj_s1(c, h);
        adj_s2(c);
    }
    const unsigned gb = (unsigned)std::min<size_t>((c->mn + 255) / 256, 2048);
    if (c->nb)
        hipLaunchKernelGGL(k2_sum_shots, dim3((unsigned)std::min<size_t>((3 * c->mn + 255) / 256, 4096)), dim3(256), 0,
                           c->stream, c->bt.grad, ns, (long)(3 * c->mn), c->grad);
    if (para(c) != 2)
        hipLaunchKernelGGL(k2_grad_para, dim3(gb), dim3(256), 0, c->stream, c->grad, model(c, 1), geo(c), para(c));
    HC(c, hipGetLastError());
    return 0;
}

// Reference-mode adjoint over the shots of the running operator (one, or nb
// batched along z): background pass storing E and F per step, then the
// reversed adjoint steps from the residual records in rec_of(c), gradient
// (g_lam, g_mu, g_rho) of every step accumulated per shot and summed into
// c->grad.  mapped: the records are first scaled per receiver by dt / rho
// and the sum is mapped to the inverted parameters (k2r_grad_map).
int adjoint_ref_core(erm2d_ctx* c, bool mapped) {
    if (ls_strips(c)) return c->fail(ERM_E_PARAM, "the reference-mode adjoint needs adjoint_source 0 (history)");
    int rc;
    const long comp = (long)c->p.lt * c->p.nrec;
    const int rg = (c->p.nrec + 255) / 256;
    const int ns = nshot_of(c);
    if ((rc = zero_state(c, false)) || (rc = zero_state(c, true))) return rc;
    if ((rc = fillf(c, grad_of(c), 3 * c->mn * ns, 0.f))) return rc;
    float* hist = hist_of(c);
    float* rec = rec_of(c);
    const long sst = c->nb ? (long)c->ncell : 0, rst = c->nb ? 2 * comp : 0;
    const F2 fb = state(c, true);
    const size_t step = 5 * c->mn;
    if (mapped && c->p.nrec > 0)
        hipLaunchKernelGGL(k2r_scale_rec, dim3((unsigned)std::min<long>((2 * comp + 255) / 256, 1024), ns), dim3(256), 0,
                           c->stream, rec, c->rec_idx, c->p.nrec, (long)c->p.lt, c->mod[1].B, c->dtr, rst);
    for (int it = 0; it < c->p.lt; ++it) {
        float* h = hist + (size_t)it * step;
        do_stress(c, false, 1, c->src, wavelet_at(c, it), h);
        do_velocity(c, false, 1, 0, it, h + 3 * c->mn);
    }
    for (int it = c->p.lt - 1; it >= 0; --it) {
        if (c->p.nrec > 0)
            hipLaunchKernelGGL(k2_inject, dim3(rg, ns), dim3(256), 0, c->stream, fb.vx, fb.vz, c->rec_idx, c->p.nrec,
                               rec, comp, (long)it * c->p.nrec, sst, rst);
        ref_step(c, hist + (size_t)it * step);
    }
    if (c->nb)
        hipLaunchKernelGGL(k2_sum_shots, dim3((unsigned)std::min<size_t>((3 * c->mn + 255) / 256, 4096)), dim3(256), 0,
                           c->stream, c->bt.grad, ns, (long)(3 * c->mn), c->grad);
    if (mapped)
        hipLaunchKernelGGL(k2r_grad_map, dim3((unsigned)std::min<size_t>((c->mn + 255) / 256, 2048)), dim3(256), 0,
                           c->stream, c->grad, model(c, 1), geo(c), para(c));
    HC(c, hipGetLastError());
    return 0;
}

int lsrtm_ready(erm2d_ctx* c) {
    if (c->cpml) return c->fail(ERM_E_PARAM, "the LSRTM operators need the damping boundary (boundary = 0)");
    HC(c, hipSetDevice(c->device));
    if (!c->mod[1].have) return c->fail(ERM_E_STATE, "background model not loaded");
    return ensure_lsrtm(c);
}
int timed_ms(erm2d_ctx* c) {
    HC(c, hipEventRecord(c->ev[1], c->stream));
    HC(c, hipEventSynchronize(c->ev[1]));
    float ms = 0;
    HC(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->last_ms = ms;
    return 0;
}
// source cells of shots [k0, k0 + ns) into the batch table (validated)
int batch_sources(erm2d_ctx* c, const int* shots, int ns) {
    std::vector<long> src(ns);
    for (int k = 0; k < ns; ++k) {
        const int i = c->p.bl + shots[2 * k], kk = c->p.bu + shots[2 * k + 1];
        if (i < c->R || i >= c->nnx - c->R || kk < c->R || kk >= c->nnz - c->R)
            return c->fail(ERM_E_PARAM, "source (%d,%d) is not an updated cell", i, kk);
        src[k] = (long)i * c->pz + kk;
    }
    HC(c, hipMemcpyAsync(c->bt.srcs, src.data(), ns * sizeof(long), hipMemcpyHostToDevice, c->stream));
    return 0;
}
void free_batch(erm2d_ctx* c) {
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    auto& b = c->bt;
    for (int q = 0; q < 8; ++q) { F(b.A[q]); F(b.B[q]); }
    for (int l = 0; l < 3; ++l) for (int q = 0; q < 4; ++q) { F(b.rA[l][q]); F(b.rB[l][q]); }
    for (auto& w : b.W) F(w);
    F(b.hist); F(b.grad); F(b.rec); F(b.srcs); F(b.strip);
    b = erm2d_ctx::Batch{};
}
}  // namespace

extern "C" {

int erm2d_born(erm2d_ctx* c, const float* dm, float* d) {
    if (!c || !dm) return ERM_E_PARAM;
    int rc;
    if ((rc = lsrtm_ready(c))) return rc;
    HC(c, hipMemcpyAsync(c->dm, dm, 3 * c->mn * 4, hipMemcpyHostToDevice, c->stream));
    HC(c, hipEventRecord(c->ev[0], c->stream));
    c->nb = 0;
    if ((rc = born_core(c)) || (rc = timed_ms(c))) return rc;
    if (d && c->p.nrec > 0) {
        HC(c, hipMemcpyAsync(d, c->rec2, (size_t)2 * c->p.lt * c->p.nrec * 4, hipMemcpyDeviceToHost, c->stream));
        HC(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int erm2d_adjoint(erm2d_ctx* c, const float* r, float* gout) {
    if (!c || !r) return ERM_E_PARAM;
    int rc;
    if ((rc = lsrtm_ready(c))) return rc;
    if (c->p.nrec > 0)
        HC(c, hipMemcpyAsync(c->rec2, r, (size_t)2 * c->p.lt * c->p.nrec * 4, hipMemcpyHostToDevice, c->stream));
    HC(c, hipEventRecord(c->ev[0], c->stream));
    c->nb = 0;
    if ((rc = adjoint_core(c)) || (rc = timed_ms(c))) return rc;
    if (gout) {
        HC(c, hipMemcpyAsync(gout, c->grad, 3 * c->mn * 4, hipMemcpyDeviceToHost, c->stream));
        HC(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int erm2d_adjoint_ref(erm2d_ctx* c, const float* r, float* gout) {
    if (!c || !r) return ERM_E_PARAM;
    int rc;
    if ((rc = lsrtm_ready(c))) return rc;
    if (c->p.nrec > 0)
        HC(c, hipMemcpyAsync(c->rec2, r, (size_t)2 * c->p.lt * c->p.nrec * 4, hipMemcpyHostToDevice, c->stream));
    HC(c, hipEventRecord(c->ev[0], c->stream));
    c->nb = 0;
    if ((rc = adjoint_ref_core(c, false)) || (rc = timed_ms(c))) return rc;
    if (gout) {
        HC(c, hipMemcpyAsync(gout, c->grad, 3 * c->mn * 4, hipMemcpyDeviceToHost, c->stream));
        HC(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int erm2d_set_batch(erm2d_ctx* c, int nb) {
    if (!c) return ERM_E_PARAM;
    if (nb < 1 || nb > 1024) return c->fail(ERM_E_PARAM, "batch must be 1..1024 shots");
    HC(c, hipSetDevice(c->device));
    HC(c, hipStreamSynchronize(c->stream));
    free_batch(c);
    if (nb == 1) return 0;
    auto& b = c->bt;
    const size_t n = c->ncell * nb;
    int rc = 0;
    for (int q = 0; q < 8 && !rc; ++q) rc = dmal(c, &b.A[q], n) || dmal(c, &b.B[q], n);
    for (int l = 0; l < c->L && !rc; ++l)
        for (int q = 0; q < 4 && !rc; ++q) rc = dmal(c, &b.rA[l][q], n) || dmal(c, &b.rB[l][q], n);
    for (int q = 0; q < 3 && !rc; ++q) rc = dmal(c, &b.W[q], n);
    if (!rc) rc = dmal(c, &b.hist, (size_t)hist_steps(c) * 5 * c->mn * nb);
    if (!rc && ls_strips(c)) {
        rc = ensure_ring(c);
        if (!rc) rc = dmal(c, &b.strip, (size_t)c->p.lt * 8 * c->xc_nr * nb);
    }
    if (!rc) rc = dmal(c, &b.grad, 3 * c->mn * nb);
    if (!rc) rc = dmal(c, &b.rec, std::max<size_t>((size_t)2 * c->p.lt * c->p.nrec * nb, 1));
    if (!rc) rc = dmal(c, &b.srcs, nb);
    if (rc) {
        free_batch(c);
        return c->fail(ERM_E_NOMEM, "batch of %d shots does not fit in device memory", nb);
    }
    b.cap = nb;
    return 0;
}

int erm2d_born_shots(erm2d_ctx* c, int nshots, const int* shots, const float* dm, float* d) {
    if (!c || nshots < 1 || !shots || !dm) return c ? c->fail(ERM_E_PARAM, "bad arguments") : ERM_E_PARAM;
    int rc;
    if ((rc = lsrtm_ready(c))) return rc;
    const size_t nd = (size_t)2 * c->p.lt * c->p.nrec;
    if (c->bt.cap < 2) {  // one shot at a time on this context
        const int sx0 = c->p.shot_x, sz0 = c->p.shot_z;
        for (int k = 0; k < nshots && !rc; ++k)
            if (!(rc = erm2d_set_shot(c, shots[2 * k], shots[2 * k + 1]))) rc = erm2d_born(c, dm, d ? d + k * nd : nullptr);
        erm2d_set_shot(c, sx0, sz0);
        return rc;
    }
    HC(c, hipMemcpyAsync(c->dm, dm, 3 * c->mn * 4, hipMemcpyHostToDevice, c->stream));
    HC(c, hipEventRecord(c->ev[0], c->stream));
    for (int k0 = 0; k0 < nshots; k0 += c->bt.cap) {
        const int ns = std::min(c->bt.cap, nshots - k0);
        if ((rc = batch_sources(c, shots + 2 * k0, ns))) return rc;
        c->nb = ns;
        rc = born_core(c);
        c->nb = 0;
        if (rc) return rc;
        if (d && nd) HC(c, hipMemcpyAsync(d + k0 * nd, c->bt.rec, nd * ns * 4, hipMemcpyDeviceToHost, c->stream));
    }
    if ((rc = timed_ms(c))) return rc;
    HC(c, hipStreamSynchronize(c->stream));
    return 0;
}

int erm2d_adjoint_shots(erm2d_ctx* c, int nshots, const int* shots, const float* r, float* gout) {
    if (!c || nshots < 1 || !shots || !r) return c ? c->fail(ERM_E_PARAM, "bad arguments") : ERM_E_PARAM;
    int rc;
    if ((rc = lsrtm_ready(c))) return rc;
    const size_t nd = (size_t)2 * c->p.lt * c->p.nrec, nm = 3 * c->mn;
    std::vector<double> acc(nm, 0.0);
    std::vector<float> part(nm);
    auto add = [&]() { for (size_t i = 0; i < nm; ++i) acc[i] += part[i]; };
    if (c->bt.cap < 2) {
        const int sx0 = c->p.shot_x, sz0 = c->p.shot_z;
        for (int k = 0; k < nshots && !rc; ++k)
            if (!(rc = erm2d_set_shot(c, shots[2 * k], shots[2 * k + 1])) && !(rc = erm2d_adjoint(c, r + k * nd, part.data())))
                add();
        erm2d_set_shot(c, sx0, sz0);
    } else {
        HC(c, hipEventRecord(c->ev[0], c->stream));
        for (int k0 = 0; k0 < nshots && !rc; k0 += c->bt.cap) {
            const int ns = std::min(c->bt.cap, nshots - k0);
            if ((rc = batch_sources(c, shots + 2 * k0, ns))) return rc;
            if (nd) HC(c, hipMemcpyAsync(c->bt.rec, r + k0 * nd, nd * ns * 4, hipMemcpyHostToDevice, c->stream));
            c->nb = ns;
            rc = adjoint_core(c);
            c->nb = 0;
            if (rc) return rc;
            HC(c, hipMemcpyAsync(part.data(), c->grad, nm * 4, hipMemcpyDeviceToHost, c->stream));
            HC(c, hipStreamSynchronize(c->stream));
            add();
        }
        if (!rc) rc = timed_ms(c);
    }
    if (!rc && gout) for (size_t i = 0; i < nm; ++i) gout[i] = (float)acc[i];
    return rc;
}

}  // extern "C"

namespace {
// Angle gathers into the device staging d of nd = 2 nang mn + 3 mn floats
// ([nang][mn] pp, [nang][mn] ps, [mn] ppc, [mn] psc, [mn] gamma), which is
// cleared first, then copied to the host outputs.
int angle_gathers_into(erm2d_ctx* c, float* d, size_t nd, int nang, float dang, float* pp, float* ps, float* pp_comp,
                       float* ps_comp, float* gamma) {
    const size_t mn = c->mn;
    int rc = fillf(c, d, nd, 0.f);
    if (rc) return rc;
    const G2 g = geo(c);
    hipLaunchKernelGGL(k2_adcig, dim3((unsigned)std::min<size_t>((mn + 255) / 256, 2048)), dim3(256), 0, c->stream,
                       excite(c), g, nang, dang, d, d + nang * mn, d + 2 * nang * mn, d + 2 * nang * mn + mn,
                       d + 2 * nang * mn + 2 * mn);
    hipError_t e = hipGetLastError();
    float* outs[5] = {pp, ps, pp_comp, ps_comp, gamma};
    const size_t off[5] = {0, nang * mn, 2 * nang * mn, 2 * nang * mn + mn, 2 * nang * mn + 2 * mn};
    const size_t len[5] = {nang * mn, nang * mn, mn, mn, mn};
    for (int a = 0; a < 5 && e == hipSuccess; ++a)
        if (outs[a]) e = hipMemcpyAsync(outs[a], d + off[a], len[a] * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->fail(ERM_E_DEVICE, "angle gathers: %s", hipGetErrorString(e));
    return 0;
}
}  // namespace

extern "C" {

int erm2d_angle_gathers(erm2d_ctx* c, int nang, float dang, float* pp, float* ps, float* pp_comp,
                        float* ps_comp, float* gamma) {
    if (!c) return ERM_E_PARAM;
    if (nang <= 0 || !(dang > 0.f) || nang * dang > 180.f) return c->fail(ERM_E_PARAM, "bad angle bins");
    HC(c, hipSetDevice(c->device));
    const size_t nd = 2 * (size_t)nang * c->mn + 3 * c->mn;
    float* d = nullptr;
    if (dmal(c, &d, nd)) return ERM_E_NOMEM;
    const int rc = angle_gathers_into(c, d, nd, nang, dang, pp, ps, pp_comp, ps_comp, gamma);
    (void)hipFree(d);
    return rc;
}

int erm2d_angle_gathers_noalloc(erm2d_ctx* c, int nang, float dang, float* pp, float* ps, float* pp_comp,
                                float* ps_comp, float* gamma) {
    if (!c) return ERM_E_PARAM;
    if (nang <= 0 || !(dang > 0.f) || nang * dang > 180.f) return c->fail(ERM_E_PARAM, "bad angle bins");
    const size_t nd = 2 * (size_t)nang * c->mn + 3 * c->mn;
    if (nd > kScratchFloats)
        return c->fail(ERM_E_SIZE, "angle gathers need %zu staging floats (2 nang nx nz + 3 nx nz); the allocation-free "
                                   "entry point holds at most %zu", nd, kScratchFloats);
    HC(c, hipSetDevice(c->device));
    float* d = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
    if (!d) return c->fail(ERM_E_STATE, "scratch address of device %d not resolved", c->device);
    return angle_gathers_into(c, d, nd, nang, dang, pp, ps, pp_comp, ps_comp, gamma);
}

int erm2d_illumination(erm2d_ctx* c, int nshots, const int* shots, float* out) {
    if (!c) return ERM_E_PARAM;
    if (nshots <= 0 || !shots || !out) return c->fail(ERM_E_PARAM, "bad arguments");
    HC(c, hipSetDevice(c->device));
    if (!c->mod[1].have) return c->fail(ERM_E_STATE, "background model not loaded");
    if (!c->illum && dmal(c, &c->illum, c->ncell)) return ERM_E_NOMEM;
    int rc = fillf(c, c->illum, c->ncell, 0.f);
    const int sx0 = c->p.shot_x, sz0 = c->p.shot_z;
    const long n = (long)c->ncell;
    for (int s = 0; s < nshots && !rc; ++s) {
        if ((rc = erm2d_set_shot(c, shots[2 * s], shots[2 * s + 1])) || (rc = zero_state(c, false))) break;
        for (int it = 0; it < c->p.lt; ++it) {
            do_stress(c, false, 1, c->src, wavelet_at(c, it), nullptr);
            do_velocity(c, false, 1, ERM2D_PASS_MODEL, it, nullptr);
            hipLaunchKernelGGL(k2_illum_acc, dim3((unsigned)std::min<long>((n + 255) / 256, 2048)), dim3(256), 0,
                               c->stream, c->fA[0], c->fA[1], c->illum, n);
        }
        HC(c, hipGetLastError());
    }
    if (!rc) rc = cut_model(c, c->illum, out);
    const int rc2 = erm2d_set_shot(c, sx0, sz0);
    return rc ? rc : rc2;
}

int erm2d_precond_diag(const float* g, float* z, void* user) {
    const erm2d_diag_precond* d = (const erm2d_diag_precond*)user;
    if (!d || !d->w || d->nx <= 0 || d->nz <= 0 || !g || !z) return ERM_E_PARAM;
    const size_t mn = (size_t)d->nx * d->nz;
    for (int ch = 0; ch < 3; ++ch)
        for (size_t i = 0; i < mn; ++i) z[ch * mn + i] = d->w[i] * g[ch * mn + i];
    return 0;
}

int erm2d_rtm_windows(const erm2d_params* p, int device, const float* vp, const float* vs, const float* rho,
                      const float* qp, const float* qs, const float* bvp, const float* bvs, const float* brho,
                      const float* bqp, const float* bqs, int nshots, const erm2d_window* win, const float* records,
                      int z2, float* img_tp, float* img_pp, float* img_ps, float* illum) {
    if (!p || !vp || !vs || !rho || !bvp || !bvs || !brho || nshots <= 0 || !win) { g_err2 = "bad arguments"; return ERM_E_PARAM; }
    if (p->nsls > 0 && (!qp || !qs || !bqp || !bqs)) { g_err2 = "qp/qs required when nsls > 0"; return ERM_E_PARAM; }
    if (p->rec_dx <= 0 || p->nx <= 0 || p->nz <= 0) { g_err2 = "bad grid or receiver spacing"; return ERM_E_PARAM; }
    const int NX = p->nx, NZ = p->nz;
    for (int s = 0; s < nshots; ++s) {
        const erm2d_window& w = win[s];
        if (w.x0 < 0 || w.width < p->rec_dx || w.x0 + w.width > NX || w.shot_x < w.x0 || w.shot_x >= w.x0 + w.width ||
            w.shot_z < 0 || w.shot_z >= NZ) {
            g_err2 = erm::format("window %d outside the model or not containing its shot", s);
            return ERM_E_PARAM;
        }
    }
    const size_t fm = (size_t)NX * NZ;
    std::vector<double> acc[4];
    for (auto& a : acc) a.assign(fm, 0.0);
    std::vector<erm2d_ctx*> ctxs;  // one context per window width
    std::vector<int> widths;
    auto release = [&]() { for (auto* c : ctxs) erm2d_destroy(c); };
    size_t rec_off = 0;
    int rc = 0;
    erm2d_ctx* cur = nullptr;
    for (int s = 0; s < nshots && !rc; ++s) {
        const erm2d_window& w = win[s];
        erm2d_ctx* c = nullptr;
        for (size_t q = 0; q < widths.size(); ++q)
            if (widths[q] == w.width) c = ctxs[q];
        const int nrec = w.width / p->rec_dx;
        if (!c) {
            erm2d_params q = *p;
            q.nx = w.width;
            q.nrec = nrec;
            q.rec_x0 = 0;
            q.shot_x = w.shot_x - w.x0;
            q.shot_z = w.shot_z;
            if ((rc = erm2d_create(&q, device, &c))) break;
            ctxs.push_back(c);
            widths.push_back(w.width);
        }
        cur = c;
        // window models: cuda_get_partly_mode_boundary (+ _z1_z2 on the background)
        const size_t wm = (size_t)w.width * NZ;
        std::vector<float> cut[10];
        const float* src[10] = {vp, vs, rho, qp, qs, bvp, bvs, brho, bqp, bqs};
        for (int a = 0; a < 10; ++a) {
            if (!src[a] || (p->nsls == 0 && (a == 3 || a == 4 || a == 8 || a == 9))) continue;
            cut[a].resize(wm);
            for (int ix = 0; ix < w.width; ++ix)
                for (int iz = 0; iz < NZ; ++iz) {
                    const int zz = (a >= 5 && z2 >= 0 && iz > z2) ? std::min(z2, NZ - 1) : iz;
                    cut[a][(size_t)ix * NZ + iz] = src[a][(size_t)(ix + w.x0) * NZ + zz];
                }
        }
        auto ptr = [&](int a) { return cut[a].empty() ? nullptr : cut[a].data(); };
        if ((rc = erm2d_set_shot(c, w.shot_x - w.x0, w.shot_z))) break;
        const float* rec = nullptr;
        std::vector<float> modelled;
        if (records) {
            rec = records + rec_off;
            rec_off += (size_t)2 * p->lt * nrec;
        } else {
            if ((rc = erm2d_load_model(c, 0, ptr(0), ptr(1), ptr(2), ptr(3), ptr(4)))) break;
            modelled.resize((size_t)2 * p->lt * nrec);
            if ((rc = erm2d_forward(c, modelled.data()))) break;
            rec = modelled.data();
        }
        if ((rc = erm2d_load_model(c, 1, ptr(5), ptr(6), ptr(7), ptr(8), ptr(9)))) break;
        std::vector<float> im[4];
        for (auto& v : im) v.assign(wm, 0.f);
        if ((rc = erm2d_rtm(c, rec, im[0].data(), im[1].data(), im[2].data()))) break;
        if (illum) {
            const int sh[2] = {w.shot_x - w.x0, w.shot_z};
            if ((rc = erm2d_illumination(c, 1, sh, im[3].data()))) break;
        }
        // cuda_sum_new_acqusition[_illum]: add the window into the full model at x0
        for (int a = 0; a < 4; ++a)
            for (size_t t = 0; t < wm; ++t) acc[a][(size_t)w.x0 * NZ + t] += im[a][t];
    }
    if (rc) {
        if (cur && !cur->err.empty()) g_err2 = cur->err;
        release();
        return rc;
    }
    release();
    float* outs[4] = {img_tp, img_pp, img_ps, illum};
    for (int a = 0; a < 4; ++a)
        if (outs[a]) for (size_t t = 0; t < fm; ++t) outs[a][t] = (float)acc[a][t];
    return 0;
}

int erm2d_last_ms(const erm2d_ctx* c, double* ms) {
    if (!c || !ms) return ERM_E_PARAM;
    *ms = c->last_ms;
    return 0;
}

}  // extern "C"

extern "C" int erm2d_lsrtm_cg(erm2d_ctx* c, int nshots, const int* shots, const float* dobs, int iters, unsigned pmask,
                              float* m, float* misfit, double* seconds, erm2d_precond_fn precond, void* user,
                              float tol) {
    if (!c || nshots <= 0 || !shots || !dobs || iters < 0 || !m || !misfit) return c ? c->fail(ERM_E_PARAM, "bad arguments") : ERM_E_PARAM;
    if (c->cpml) return c->fail(ERM_E_PARAM, "the LSRTM operators need the damping boundary (boundary = 0)");
    const size_t nm = 3 * c->mn, nd = (size_t)2 * c->p.lt * c->p.nrec;
    const int sx0 = c->p.shot_x, sz0 = c->p.shot_z;
    auto t0 = std::chrono::steady_clock::now();
    auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    auto mask = [&](std::vector<float>& v) {
        for (int a = 0; a < 3; ++a)
            if (!(pmask & (1u << a))) std::fill(v.begin() + a * c->mn, v.begin() + (a + 1) * c->mn, 0.f);
    };
    auto dot = [](const float* a, const float* b, size_t n) { double s = 0; for (size_t i = 0; i < n; ++i) s += (double)a[i] * b[i]; return s; };
    std::vector<float> r((size_t)nshots * nd), q((size_t)nshots * nd), s(nm), z(nm), pvec(nm);
    std::vector<float> mv(m, m + nm);
    mask(mv);
    // Shots are spread over the context and its workers; each worker runs its
    // shots on its own stream from its own host thread, so small shots overlap.
    std::vector<erm2d_ctx*> pool{c};
    for (auto* w : c->workers) pool.push_back(w);
    const int nw = std::min((int)pool.size(), nshots);
    std::vector<std::vector<float>> part(nw, std::vector<float>(nm)), tmp(nw, std::vector<float>(nm));
    auto parallel = [&](auto&& job) -> int {
        std::vector<int> rcs(nw, 0);
        std::vector<std::thread> th;
        for (int w = 1; w < nw; ++w)
            th.emplace_back([&, w]() {
                for (int k = w; k < nshots && !rcs[w]; k += nw) rcs[w] = job(pool[w], w, k);
            });
        for (int k = 0; k < nshots && !rcs[0]; k += nw) rcs[0] = job(pool[0], 0, k);
        for (auto& t : th) t.join();
        for (int w = 0; w < nw; ++w)
            if (rcs[w]) return w ? c->fail(rcs[w], "worker %d: %s", w, pool[w]->err.c_str()) : rcs[w];
        return 0;
    };
    // L over all shots: batched launches on this context (erm2d_set_batch),
    // or one shot per launch spread over the workers
    const bool batched = c->bt.cap > 1;
    auto applyL = [&](const std::vector<float>& x, std::vector<float>& out) -> int {
        if (batched) return erm2d_born_shots(c, nshots, shots, x.data(), out.data());
        return parallel([&](erm2d_ctx* cw, int, int k) {
            int rc = erm2d_set_shot(cw, shots[2 * k], shots[2 * k + 1]);
            if (!rc) rc = erm2d_born(cw, x.data(), out.data() + (size_t)k * nd);
            return rc;
        });
    };
    auto applyLT = [&](const std::vector<float>& y, std::vector<float>& out) -> int {
        if (batched) {
            int rc2 = erm2d_adjoint_shots(c, nshots, shots, y.data(), out.data());
            if (!rc2) mask(out);
            return rc2;
        }
        for (auto& pp : part) std::fill(pp.begin(), pp.end(), 0.f);
        int rc = parallel([&](erm2d_ctx* cw, int w, int k) {
            int rc2 = erm2d_set_shot(cw, shots[2 * k], shots[2 * k + 1]);
            if (!rc2) rc2 = erm2d_adjoint(cw, y.data() + (size_t)k * nd, tmp[w].data());
            if (!rc2) for (size_t i = 0; i < nm; ++i) part[w][i] += tmp[w][i];
            return rc2;
        });
        if (rc) return rc;
        std::fill(out.begin(), out.end(), 0.f);
        for (int w = 0; w < nw; ++w)
            for (size_t i = 0; i < nm; ++i) out[i] += part[w][i];
        mask(out);
        return 0;
    };
    auto pre = [&](const std::vector<float>& g, std::vector<float>& zo) -> int {
        if (!precond) { zo = g; return 0; }
        int rc = precond(g.data(), zo.data(), user);
        if (rc) return c->fail(ERM_E_PARAM, "preconditioner returned %d", rc);
        mask(zo);
        return 0;
    };
    const double dnorm = std::sqrt(dot(dobs, dobs, (size_t)nshots * nd));
    if (dnorm == 0) return c->fail(ERM_E_PARAM, "observed data are zero");
    int rc;
    if ((rc = applyL(mv, q))) return rc;
    for (size_t i = 0; i < r.size(); ++i) r[i] = dobs[i] - q[i];
    misfit[0] = (float)(std::sqrt(dot(r.data(), r.data(), r.size())) / dnorm);
    if (seconds) seconds[0] = now();
    auto notify = [&](int it) -> int {
        if (!c->monitor) return 0;
        int rc2 = c->monitor(it, mv.data(), s.data(), misfit[it], c->monitor_user);
        return rc2 ? c->fail(ERM_E_PARAM, "monitor returned %d", rc2) : 0;
    };
    if ((rc = applyLT(r, s)) || (rc = notify(0)) || (rc = pre(s, z))) return rc;
    int done = 0;
    if (!precond) {
        // CGLS (Paige & Saunders 1982)
        pvec = z;
        double gamma = dot(s.data(), z.data(), nm);
        for (int it = 1; it <= iters; ++it) {
            if (misfit[it - 1] <= tol || gamma <= 0) break;
            if ((rc = applyL(pvec, q))) return rc;
            const double qq = dot(q.data(), q.data(), q.size());
            if (qq <= 0) break;
            const double alpha = gamma / qq;
            for (size_t i = 0; i < nm; ++i) mv[i] += (float)(alpha * pvec[i]);
            for (size_t i = 0; i < r.size(); ++i) r[i] -= (float)(alpha * q[i]);
            misfit[it] = (float)(std::sqrt(dot(r.data(), r.data(), r.size())) / dnorm);
            if (seconds) seconds[it] = now();
            done = it;
            if (it == iters || misfit[it] <= tol) break;
            if ((rc = applyLT(r, s)) || (rc = notify(it)) || (rc = pre(s, z))) return rc;
            const double gnew = dot(s.data(), z.data(), nm);
            const double beta = gnew / gamma;
            for (size_t i = 0; i < nm; ++i) pvec[i] = (float)(z[i] + beta * pvec[i]);
            gamma = gnew;
        }
    } else {
        // Flexible variant for a nonlinear preconditioner (GCR-type, as in
        // Notay 2000 / Eisenstat-Elman-Schultz 1983 applied to the normal
        // equations): the new direction p = z is made L^T L-orthogonal to the
        // last kMem directions through their data-space images q_j = L p_j
        // (kept mutually orthogonal), then the step is the exact minimiser of
        // |r - a L p|.  The misfit cannot increase whatever z is, and with
        // z = L^T r the iterates are those of CGLS.
        constexpr int kMem = 4;
        std::vector<std::vector<float>> P, Q;
        std::vector<double> QQ;
        for (int it = 1; it <= iters; ++it) {
            if (misfit[it - 1] <= tol) break;
            std::vector<float> pn(z), qn(r.size());
            if ((rc = applyL(pn, qn))) return rc;
            for (size_t j = 0; j < Q.size(); ++j) {
                const double b = dot(qn.data(), Q[j].data(), qn.size()) / QQ[j];
                for (size_t i = 0; i < nm; ++i) pn[i] -= (float)(b * P[j][i]);
                for (size_t i = 0; i < qn.size(); ++i) qn[i] -= (float)(b * Q[j][i]);
            }
            const double qq = dot(qn.data(), qn.data(), qn.size());
            if (!(qq > 0)) break;
            const double alpha = dot(r.data(), qn.data(), r.size()) / qq;
            for (size_t i = 0; i < nm; ++i) mv[i] += (float)(alpha * pn[i]);
            for (size_t i = 0; i < r.size(); ++i) r[i] -= (float)(alpha * qn[i]);
            misfit[it] = (float)(std::sqrt(dot(r.data(), r.data(), r.size())) / dnorm);
            if (seconds) seconds[it] = now();
            done = it;
            if ((int)P.size() == kMem) { P.erase(P.begin()); Q.erase(Q.begin()); QQ.erase(QQ.begin()); }
            P.push_back(std::move(pn));
            Q.push_back(std::move(qn));
            QQ.push_back(qq);
            if (it == iters || misfit[it] <= tol) break;
            if ((rc = applyLT(r, s)) || (rc = notify(it)) || (rc = pre(s, z))) return rc;
        }
    }
    std::copy(mv.begin(), mv.end(), m);
    erm2d_set_shot(c, sx0, sz0);
    return done;
}
