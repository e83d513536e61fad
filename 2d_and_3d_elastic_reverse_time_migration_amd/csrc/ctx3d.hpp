// ctx3d.hpp - the 3D context shared by the one-shot driver (rtm3d.hip) and the
// batched driver (batch3d.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels3d.hpp"

using namespace erm;

struct erm_ctx {
    erm_params p;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;               // second stream of the region launches (region_streams)
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    int R = 6;
    int nnx = 0, nny = 0, nnz = 0, px = 0;
    long pxy = 0;
    size_t ncell = 0;   // allocated floats per volume (nnz*pxy)
    Coef coef{};
    float cx = 0, cy = 0, cz = 0;
    // device buffers
    float* f[13] = {nullptr};
    float* M[2] = {nullptr, nullptr};
    float* MU[2] = {nullptr, nullptr};
    float* BU[2] = {nullptr, nullptr};
    bool have_model[2] = {false, false};
    float* hxyz[3] = {nullptr, nullptr, nullptr};
    float* wavelet = nullptr;
    int wl = 0;
    std::vector<float> wavelet_h;
    float* exbuf = nullptr;     // excitation state and images (layout: Excite); the views below point into it
    int* ex_time = nullptr;
    float* ex[5] = {nullptr};   // amp tp vxp vyp vzp
    float* img[3] = {nullptr};  // tp pp ps
    float* poyn[6] = {nullptr}; // receiver Poynting rp (x y z), rs (x y z) at the imaging step (p.poynting)
    float* rec = nullptr;       // [3][lt][nr]
    long* rec_idx = nullptr;
    int nr = 0;
    bool have_record = false;
    int sxr = 0, syr = 0, szr = 0;  // shot position in model coordinates (sx_real, ...)
    long src = -1;                  // padded index of the source
    int tx = 64, ty = 8, kchunk = 32, xcd_map = 0;
    // 32-wide tiles start at the first updated column x = R, and every
    // volume's base is shifted by xoff floats inside its allocation so that
    // this column starts a 128-byte line in every row (64-wide tiles: 0, 0)
    int x0 = 0;
    size_t xoff = 0;
    // CPML (boundary == ERM_BOUNDARY_CPML)
    bool cpml = false;
    float* psi_buf = nullptr;           // C-PML state: 21 memory variables, then coefficients (layout: Pml)
    int slab_lo[3] = {0, 0, 0}, slab_hi[3] = {0, 0, 0}, slab_w[3] = {0, 0, 0};
    Pml pml{};
    // region split of the CPML launches: tiles are grouped by the set of axes
    // whose memory-variable slabs they touch (bit 0 x, 1 y, 2 z); each group
    // runs the kernel variant that updates only those memory variables
    // (group 0: the plain kernel).  Built for the current ty/kchunk.
    int* tiles[8] = {nullptr};
    int ntl[8] = {0};
    int split_tx = 0, split_ty = 0, split_kchunk = 0;
    bool split_built = false;
    bool region_split = true;
    // source-wavefield reconstruction from boundary strips
    float* fr[13] = {nullptr};       // reconstructed source wavefield
    int* strip_idx = nullptr;        // padded indices of the face slabs around the model region
    int n_strip = 0;
    float* strips = nullptr;         // [steps+1][13][n_strip]: sigma+tp (7) after the stress update, v+vP (6) after the velocity update
    int strip_steps = 0;
    bool have_strips = false;
    float* wavelet_neg = nullptr;    // -wavelet, for the time-reversed source
    float* ximg[3] = {nullptr};      // cross-correlation images tp, pp, ps (padded grid)
    float* hist = nullptr;           // decimated source history: [slots][4][nx*ny*nz] tp, vxp, vyp, vzp of the model region
    size_t hist_n = 0;               // floats allocated
    int hist_decim = 0;              // steps per slot (0: none)
    bool have_hist = false;
    float* decbuf = nullptr;         // one divergence / curl volume (allocated on first use)
    double vmax_loaded = 0.0;
    double pml_extra_stress = 0.0, pml_extra_velocity = 0.0;  // avg extra bytes per cell-update
    std::string err;
    // timing
    hipEvent_t pass_ev[2] = {nullptr, nullptr};
    bool profiling = false;
    std::vector<hipEvent_t> kev;  // 3 per step when profiling
    double last_pass_ms = 0, last_stress_ms = 0, last_velocity_ms = 0;
    long last_stress_n = 0, last_velocity_n = 0;

    int fail(int code, const char* fmt, ...) {
        va_list ap;
        va_start(ap, fmt);
        err = vformat(fmt, ap);
        va_end(ap);
        return code;
    }
};

#define HIPC(ctx, call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return (ctx)->fail(ERM_E_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                               __FILE__, __LINE__);                                              \
    } while (0)

// The 3D kernels address every array through 32-bit byte offsets from the
// first plane a workgroup reads: its z sweep of kchunk planes plus R halo
// planes on either side must span less than 4 GiB.
inline bool sweep_fits(const erm_ctx* c, int kchunk) {
    return (unsigned long long)(kchunk + 2 * c->R) * (unsigned long long)c->pxy * 4ull < (1ull << 32);
}
// the largest z sweep that fits (0: not even one plane does)
inline int max_sweep(const erm_ctx* c) {
    const long long k = (long long)(((1ull << 32) - 1) / (4ull * (unsigned long long)c->pxy)) - 2 * c->R;
    return k < 1 ? 0 : (int)std::min<long long>(k, 1 << 30);
}

// First chunk boundary z0 of the z sweeps (chunk_span).  With C-PML, a chunk
// that holds slab planes runs the C-PML kernel over all its planes, so z0 is
// the shift in (R - kchunk, R] that leaves the fewest non-slab planes in such
// chunks (ties keep R); without C-PML it is R.
inline int chunk_origin(const erm_ctx* c) {
    const int R = c->R, kc = c->kchunk, top = c->nnz - R;
    if (!c->cpml || kc <= 1) return R;
    const int in0 = c->slab_lo[2] + 1, in1 = c->nnz - c->slab_hi[2] - 2;  // planes outside both z slabs
    int best = R;
    long best_cost = -1;
    for (int z0 = R; z0 > R - kc; --z0) {
        long cost = 0;
        for (int a = z0; a < top; a += kc) {
            const int kb = std::max(R, a), ke = std::min(a + kc, top);
            const int lo = std::max(kb, in0), hi = std::min(ke - 1, in1);
            const int n_in = hi >= lo ? hi - lo + 1 : 0;
            if (n_in < ke - kb) cost += n_in;
        }
        if (best_cost < 0 || cost < best_cost) { best_cost = cost; best = z0; }
    }
    return best;
}

inline Geo make_geo(const erm_ctx* c) {
    Geo g;
    g.nnx = c->nnx; g.nny = c->nny; g.nnz = c->nnz;
    g.px = c->px; g.pxy = c->pxy;
    g.cx = c->cx; g.cy = c->cy; g.cz = c->cz;
    g.hx = c->hxyz[0]; g.hy = c->hxyz[1]; g.hz = c->hxyz[2];
    g.kchunk = c->kchunk;
    g.z0 = chunk_origin(c);
    g.x0 = c->x0;
    g.ntx = ((c->x0 ? c->nnx - 2 * c->R : c->nnx) + c->tx - 1) / c->tx;
    g.nty = (c->nny + c->ty - 1) / c->ty;
    const int nch = (c->nnz - c->R - g.z0 + c->kchunk - 1) / c->kchunk;
    g.ntiles = g.ntx * g.nty * nch;
    g.xcd_map = c->xcd_map == 1 ? 1 : 0;  // mode 2 reorders only the region-split tile lists
    g.per_xcd = (g.ntiles + 7) / 8;
    return g;
}

inline Wave make_wave(const erm_ctx* c) {
    Wave w;
    float* const* f = c->f;
    w.vx = f[0]; w.vy = f[1]; w.vz = f[2];
    w.txx = f[3]; w.tyy = f[4]; w.tzz = f[5]; w.txy = f[6]; w.txz = f[7]; w.tyz = f[8];
    w.tp = f[9]; w.vxp = f[10]; w.vyp = f[11]; w.vzp = f[12];
    return w;
}

inline Excite make_excite(const erm_ctx* c) {
    Excite e;
    e.base = c->exbuf;
    e.n = (long)c->ncell;
    return e;
}

// direct-wave mute of the modelled record when p.cut_direct_wave > 0 and
// p.vsp == 0 (defined in rtm3d.hip, used by the one-shot and batched model passes)
int mute_direct(erm_ctx* c);
int mute_cut_samples(const erm_ctx* c, double vp_src, std::vector<int>& cut);
