// batch3d.hip - batched multi-shot 3D RTM (include/erm_batch.h).
//
// The reference runs one shot per pass (3D_forward_or_back_together.cu:9-453
// per shot; per-shot windows and stacking in the 2D code,
// viscoelastic_equation.cu:2606-2754).  Here the shots of a batch step
// together: per time step one stress launch and one velocity launch cover the
// tiles of every window (k_stress_batch / k_velocity_batch over a block
// table), one launch serves all receivers, and the step is captured once in a
// hipGraph and replayed with the step index in device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <utility>
#include <string>
#include <vector>

#include "../../include/erm_batch.h"
#include "common.hpp"
#include "ctx3d.hpp"
#include "kernels3d.hpp"

struct erm_batch {
    erm_params p;
    int device = 0;
    hipStream_t stream = nullptr;
    int R = 4;
    int cus = 256;
    std::vector<float> model[2][3];  // full-model host copies (vp vs rho), disk order
    bool have_model[2] = {false, false};
    std::vector<erm_shot_window> win;
    std::vector<erm_ctx*> shots;
    std::vector<char> have_rec;
    erm_batch_progress_fn progress = nullptr;
    void* progress_user = nullptr;
    // device tables
    ShotStep* d_steps[2] = {nullptr, nullptr};
    // block tables per C-PML axis group (see build_split in rtm3d.hip): group m
    // runs the kernel variant that updates the memory variables of the axes in m
    int2* d_table[8] = {nullptr};
    int total[8] = {0};
    int total_blocks = 0;  // all groups
    RecShot* d_rec = nullptr;
    int* d_shot_of = nullptr;
    int total_rec = 0;
    int* d_it = nullptr;
    float* d_stack = nullptr;
    // launch shape
    int tx = 64, ty = 8, kchunk = 8, grid = 0;  // tiles 64 x {4, 8} or (radius 4) 32 x {8, 16}
    bool use_graph = true;
    bool tables_dirty = true;
    // timing
    hipEvent_t ev[2] = {nullptr, nullptr};
    double ms[3] = {0, 0, 0};
    int steps[3] = {0, 0, 0};
    double cells_per_step = 0;
    std::string err;
    int fail(int code, const char* fmt, ...) {
        va_list ap;
        va_start(ap, fmt);
        err = erm::vformat(fmt, ap);
        va_end(ap);
        return code;
    }
};

#define BHC(b, call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) return (b)->fail(ERM_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

namespace {

thread_local std::string g_berr;

void free_tables(erm_batch* b) {
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    F(b->d_steps[0]); F(b->d_steps[1]); F(b->d_rec); F(b->d_shot_of);
    for (auto& t : b->d_table) { F(t); t = nullptr; }
    b->d_steps[0] = b->d_steps[1] = nullptr;
    b->d_rec = nullptr;
    b->d_shot_of = nullptr;
}

void free_shots(erm_batch* b) {
    for (auto* c : b->shots) erm_destroy(c);
    b->shots.clear();
    b->have_rec.clear();
}

int load_window(erm_batch* b, int s, int which) {
    const erm_shot_window& w = b->win[s];
    const erm_params& p = b->p;
    std::vector<float> cut[3];
    for (int a = 0; a < 3; ++a) {
        cut[a].resize((size_t)w.wx * w.wy * p.nz);
        const float* src = b->model[which][a].data();
        for (int i = 0; i < w.wx; ++i)
            for (int j = 0; j < w.wy; ++j)
                std::memcpy(&cut[a][((size_t)i * w.wy + j) * p.nz], &src[((size_t)(w.x0 + i) * p.ny + (w.y0 + j)) * p.nz],
                            (size_t)p.nz * sizeof(float));
    }
    int rc = erm_load_model(b->shots[s], which, cut[0].data(), cut[1].data(), cut[2].data());
    if (rc) return b->fail(rc, "shot %d: %s", s, erm_last_error(b->shots[s]));
    return 0;
}

// XCD y-band order of a block table (xcd_map 2; the one-shot tile lists use
// the same order): workgroup b runs on XCD b % 8, each XCD with its own L2.
// The table, sorted (shot, y, z chunk, x), is cut into 8 chunks of equal
// length; XCD x takes chunk x at positions x, x + 8, ..., in (shot, z chunk,
// y, x) order, so the workgroups an XCD holds at once form a compact patch of
// one window's z layer whose x and y halos its own L2 serves.  Positions past
// a shorter chunk hold a sentinel (shot 0, tile id past its tiles) that the
// kernels skip.
std::vector<int2> band_order(const std::vector<int2>& t, const std::vector<Geo>& geo) {
    const int n = (int)t.size();
    auto key = [&](const int2& e, bool ymajor) {
        const Geo& g = geo[e.x];
        const long l = e.y & kTileIdMask;
        const long bx = l % g.ntx, by = (l / g.ntx) % g.nty, bz = l / ((long)g.ntx * g.nty);
        const long in = ymajor ? (by * g.ntiles + bz) * g.ntx + bx : (bz * g.nty + by) * g.ntx + bx;
        return std::make_pair(e.x, in);
    };
    std::vector<int2> s = t;
    std::sort(s.begin(), s.end(), [&](const int2& a, const int2& b) { return key(a, true) < key(b, true); });
    const int per = (n + 7) / 8;
    std::vector<int2> out((size_t)8 * per, make_int2(0, geo[0].ntiles));
    for (int x = 0; x < 8; ++x) {
        const int b0 = std::min(n, x * per), b1 = std::min(n, b0 + per);
        std::sort(s.begin() + b0, s.begin() + b1,
                  [&](const int2& a, const int2& b) { return key(a, false) < key(b, false); });
        for (int j = 0; j < b1 - b0; ++j) out[(size_t)8 * j + x] = s[b0 + j];
    }
    return out;
}

// Device tables for the current shots and launch shape.
int build_tables(erm_batch* b) {
    if (!b->tables_dirty) return 0;
    BHC(b, hipSetDevice(b->device));
    const int ns = (int)b->shots.size();
    std::vector<ShotStep> st[2];
    std::vector<int2> table[8];
    std::vector<RecShot> rs;
    std::vector<int> shot_of;
    std::vector<Geo> geos;
    double cells = 0;
    for (int s = 0; s < ns; ++s) {
        erm_ctx* c = b->shots[s];
        c->tx = b->tx;
        c->ty = b->ty;
        c->kchunk = std::max(1, std::min(b->kchunk, c->nnz - 2 * c->R));
        if (!sweep_fits(c, c->kchunk))
            return b->fail(ERM_E_PARAM, "zchunk %d too large for shot %d's %ld-cell planes (at most %d: 32-bit offsets)",
                           c->kchunk, s, c->pxy, max_sweep(c));
        c->xcd_map = c->xcd_map == 2 ? 2 : 0;  // the table ids are plain tile ids (make_geo maps 2 to 0)
        const Geo g = make_geo(c);
        if (g.ntiles > kTileIdMask) return b->fail(ERM_E_SIZE, "shot %d: too many tiles for the block table", s);
        for (int which = 0; which < 2; ++which) {
            ShotStep d{};
            d.w = make_wave(c);
            d.ex = make_excite(c);
            d.P = c->pml;
            d.g = g;
            d.M = c->M[which];
            d.MU = c->MU[which];
            d.B = c->BU[which];
            d.src_k = c->src < 0 ? -1 : (int)(c->src / c->pxy);
            d.src_col = c->src < 0 ? -1 : (int)(c->src % c->pxy);
            st[which].push_back(d);
        }
        // same classification as the one-shot region split: plain kernel where no
        // updated cell of the tile lies in a memory-variable slab
        auto inside = [](int a, int z, int n, int lo, int hi) { return a > lo && z < n - hi - 1; };
        const int R = c->R;
        for (int l = 0; l < g.ntiles; ++l) {
            const int bx = l % g.ntx, by = (l / g.ntx) % g.nty, bz = l / (g.ntx * g.nty);
            const int i0 = std::max(g.x0 + bx * b->tx, R), i1 = std::min(g.x0 + bx * b->tx + b->tx, c->nnx - R) - 1;
            const int j0 = std::max(by * c->ty, R), j1 = std::min(by * c->ty + c->ty, c->nny - R) - 1;
            const ZSpan zs = chunk_span(g, R, bz);
            const int k0 = zs.kb, k1 = zs.ke - 1;
            if (i0 > i1 || j0 > j1 || k0 > k1) continue;
            int m = 0, axes = 0;
            if (c->cpml) {
                m = (inside(i0, i1, c->nnx, c->slab_lo[0], c->slab_hi[0]) ? 0 : 1) |
                    (inside(j0, j1, c->nny, c->slab_lo[1], c->slab_hi[1]) ? 0 : 2) |
                    (inside(k0, k1, c->nnz, c->slab_lo[2], c->slab_hi[2]) ? 0 : 4);
                axes = c->region_split ? m : 7;
                if (!c->region_split || m != 0) m = 7;  // two groups: plain and all-axes C-PML
            }
            // C-PML entries carry the axes whose slabs the tile touches (bits
            // 28..30): the kernel skips the memory variables of the others
            table[m].push_back(make_int2(s, l | (m == 7 ? axes << 28 : 0)));
        }
        geos.push_back(g);
        RecShot r{};
        r.vx = c->f[0]; r.vy = c->f[1]; r.vz = c->f[2];
        r.idx = c->rec_idx;
        r.rec = c->rec;
        r.nr = c->nr;
        r.first = (int)shot_of.size();
        r.comp = (long)c->p.lt * c->nr;
        rs.push_back(r);
        for (int q = 0; q < c->nr; ++q) shot_of.push_back(s);
        cells += (double)(c->nnx - 2 * c->R) * (c->nny - 2 * c->R) * (c->nnz - 2 * c->R);
    }
    if (b->p.xcd_map == 2 && ns > 0)
        for (auto& t : table)
            if (!t.empty()) t = band_order(t, geos);
    free_tables(b);
    for (int which = 0; which < 2; ++which) {
        BHC(b, hipMalloc(&b->d_steps[which], std::max<size_t>(1, st[which].size()) * sizeof(ShotStep)));
        if (!st[which].empty())
            BHC(b, hipMemcpy(b->d_steps[which], st[which].data(), st[which].size() * sizeof(ShotStep), hipMemcpyHostToDevice));
    }
    b->total_blocks = 0;
    for (int m = 0; m < 8; ++m) {
        b->total[m] = (int)table[m].size();
        b->total_blocks += b->total[m];
        if (table[m].empty()) continue;
        BHC(b, hipMalloc(&b->d_table[m], table[m].size() * sizeof(int2)));
        BHC(b, hipMemcpy(b->d_table[m], table[m].data(), table[m].size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    BHC(b, hipMalloc(&b->d_rec, std::max<size_t>(1, rs.size()) * sizeof(RecShot)));
    if (!rs.empty()) BHC(b, hipMemcpy(b->d_rec, rs.data(), rs.size() * sizeof(RecShot), hipMemcpyHostToDevice));
    BHC(b, hipMalloc(&b->d_shot_of, std::max<size_t>(1, shot_of.size()) * sizeof(int)));
    if (!shot_of.empty()) BHC(b, hipMemcpy(b->d_shot_of, shot_of.data(), shot_of.size() * sizeof(int), hipMemcpyHostToDevice));

    b->total_rec = (int)shot_of.size();
    b->cells_per_step = cells;
    b->tables_dirty = false;
    return 0;
}

// Group 0 (tiles without C-PML cells) runs the plain kernel, group 7 the
// all-axes C-PML kernel; a persistent grid is split between them in
// proportion to their sizes.
inline int group_grid(const erm_batch* b, int m) {
    if (b->grid <= 0 || b->grid >= b->total_blocks) return b->total[m];
    return b->total[m] ? std::max(1, (int)((long)b->grid * b->total[m] / b->total_blocks)) : 0;
}
template <int R, int TX, int TY, int AX>
void stress_group(erm_batch* b, int which, int source_on, const Coef& cf, const float* wavelet, int wl) {
    const int gr = group_grid(b, AX);
    if (gr == b->total[AX] && gr)
        hipLaunchKernelGGL((k_stress_batch<R, TX, TY, AX, false>), dim3(gr), dim3(TX, TY), 0, b->stream, b->d_steps[which],
                           b->d_table[AX], b->total[AX], cf, wavelet, wl, (const int*)b->d_it, source_on);
    else if (gr)
        hipLaunchKernelGGL((k_stress_batch<R, TX, TY, AX, true>), dim3(gr), dim3(TX, TY), 0, b->stream, b->d_steps[which],
                           b->d_table[AX], b->total[AX], cf, wavelet, wl, (const int*)b->d_it, source_on);
}
template <int R, int TX, int TY, int MODE, int AX>
void velocity_group(erm_batch* b, int which, const Coef& cf) {
    const int gr = group_grid(b, AX);
    if (gr == b->total[AX] && gr)
        hipLaunchKernelGGL((k_velocity_batch<R, TX, TY, MODE, AX, false>), dim3(gr), dim3(TX, TY), 0, b->stream,
                           b->d_steps[which], b->d_table[AX], b->total[AX], cf, (const int*)b->d_it);
    else if (gr)
        hipLaunchKernelGGL((k_velocity_batch<R, TX, TY, MODE, AX, true>), dim3(gr), dim3(TX, TY), 0, b->stream,
                           b->d_steps[which], b->d_table[AX], b->total[AX], cf, (const int*)b->d_it);
}
template <int R, int TX, int TY, bool CPML>
void launch_stress(erm_batch* b, int which, int source_on, const Coef& cf, const float* wavelet, int wl) {
    stress_group<R, TX, TY, 0>(b, which, source_on, cf, wavelet, wl);
    if constexpr (CPML) {
        stress_group<R, TX, TY, 7>(b, which, source_on, cf, wavelet, wl);
    }
}
template <int R, int TX, int TY, int MODE, bool CPML>
void launch_velocity(erm_batch* b, int which, const Coef& cf) {
    velocity_group<R, TX, TY, MODE, 0>(b, which, cf);
    if constexpr (CPML) {
        velocity_group<R, TX, TY, MODE, 7>(b, which, cf);
    }
}
template <int R, int TX, int TY, bool CPML>
void launch_step_rt(erm_batch* b, int pass) {
    erm_ctx* c0 = b->shots[0];
    const int which = pass == ERM_PASS_MODEL ? 0 : 1;
    if (pass == ERM_PASS_RECEIVER && b->total_rec > 0)
        hipLaunchKernelGGL(k_receivers_batch, dim3((b->total_rec + 255) / 256), dim3(256), 0, b->stream, b->d_rec,
                           b->d_shot_of, b->total_rec, (const int*)b->d_it, 1);
    launch_stress<R, TX, TY, CPML>(b, which, pass != ERM_PASS_RECEIVER, c0->coef, c0->wavelet, c0->wl);
    if (pass == ERM_PASS_SOURCE) launch_velocity<R, TX, TY, MODE_EXCITE, CPML>(b, which, c0->coef);
    else if (pass == ERM_PASS_RECEIVER) launch_velocity<R, TX, TY, MODE_IMAGE, CPML>(b, which, c0->coef);
    else launch_velocity<R, TX, TY, MODE_PLAIN, CPML>(b, which, c0->coef);
    if (pass == ERM_PASS_MODEL && b->total_rec > 0)
        hipLaunchKernelGGL(k_receivers_batch, dim3((b->total_rec + 255) / 256), dim3(256), 0, b->stream, b->d_rec,
                           b->d_shot_of, b->total_rec, (const int*)b->d_it, 0);
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, b->stream, b->d_it, pass == ERM_PASS_RECEIVER ? -1 : 1);
}
void launch_step(erm_batch* b, int pass) {
    const bool cpml = b->shots[0]->cpml;
    if (b->R == 4) {
        if (b->tx == 32) {
            if (b->ty == 8) { if (cpml) launch_step_rt<4, 32, 8, true>(b, pass); else launch_step_rt<4, 32, 8, false>(b, pass); }
            else { if (cpml) launch_step_rt<4, 32, 16, true>(b, pass); else launch_step_rt<4, 32, 16, false>(b, pass); }
        } else if (b->ty == 4) { if (cpml) launch_step_rt<4, 64, 4, true>(b, pass); else launch_step_rt<4, 64, 4, false>(b, pass); }
        else { if (cpml) launch_step_rt<4, 64, 8, true>(b, pass); else launch_step_rt<4, 64, 8, false>(b, pass); }
    } else {
        if (b->ty == 4) { if (cpml) launch_step_rt<6, 64, 4, true>(b, pass); else launch_step_rt<6, 64, 4, false>(b, pass); }
        else { if (cpml) launch_step_rt<6, 64, 8, true>(b, pass); else launch_step_rt<6, 64, 8, false>(b, pass); }
    }
}

// Run nsteps of `pass` for all shots starting at it_begin (receiver pass counts down).
int run_steps(erm_batch* b, int pass, int it_begin, int nsteps, bool graph, float* ms_out) {
    int rc;
    if ((rc = build_tables(b))) return rc;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, b->stream, b->d_it, 1L, it_begin);
    BHC(b, hipGetLastError());
    BHC(b, hipEventRecord(b->ev[0], b->stream));
    if (graph && nsteps > 0) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        BHC(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        launch_step(b, pass);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (e != hipSuccess) return b->fail(ERM_E_DEVICE, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e != hipSuccess) { (void)hipGraphDestroy(g); return b->fail(ERM_E_DEVICE, "graph instantiate: %s", hipGetErrorString(e)); }
        for (int s = 0; s < nsteps && e == hipSuccess; ++s) e = hipGraphLaunch(ge, b->stream);
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return b->fail(ERM_E_DEVICE, "graph launch: %s", hipGetErrorString(e));
    } else {
        for (int s = 0; s < nsteps; ++s) launch_step(b, pass);
    }
    BHC(b, hipGetLastError());
    BHC(b, hipEventRecord(b->ev[1], b->stream));
    BHC(b, hipEventSynchronize(b->ev[1]));
    float ms = 0;
    BHC(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    if (ms_out) *ms_out = ms;
    return 0;
}

int zero_shots(erm_batch* b, bool wave, bool excite, bool images) {
    for (size_t s = 0; s < b->shots.size(); ++s) {
        erm_ctx* c = b->shots[s];
        int rc = 0;
        if (wave && (rc = erm_zero_wavefield(c))) return b->fail(rc, "shot %zu: %s", s, erm_last_error(c));
        if (excite && (rc = erm_zero_excitation(c))) return b->fail(rc, "shot %zu: %s", s, erm_last_error(c));
        if (images)
            for (auto* im : c->img)
                hipLaunchKernelGGL(k_fill_f32, dim3(1024), dim3(256), 0, c->stream, im, (long)c->ncell, 0.f);
        BHC(b, hipStreamSynchronize(c->stream));
    }
    return 0;
}

}  // namespace

extern "C" {

const char* erm_batch_last_error(const erm_batch* b) { return b ? b->err.c_str() : g_berr.c_str(); }

int erm_batch_create(const erm_params* p, int device, erm_batch** out) {
    if (!p || !out) { g_berr = "null argument"; return ERM_E_PARAM; }
    *out = nullptr;
    const int R = p->fd_order == 8 ? 4 : p->fd_order == 12 ? 6 : -1;
    if (R < 0) { g_berr = "fd_order must be 8 or 12"; return ERM_E_PARAM; }
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0 || p->lt <= 0) { g_berr = "bad grid"; return ERM_E_PARAM; }
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) { g_berr = "no HIP device"; return ERM_E_DEVICE; }
    if (device < 0 || device >= nd) { g_berr = "device out of range"; return ERM_E_PARAM; }
    erm_batch* b = new erm_batch();
    b->p = *p;
    if (b->p.lt_forward < 0) b->p.lt_forward = 2 * b->p.lt / 3;
    b->device = device;
    b->R = R;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess ||
        hipMalloc(&b->d_it, sizeof(int)) != hipSuccess) {
        g_berr = "device setup failed";
        erm_batch_destroy(b);
        return ERM_E_DEVICE;
    }
    b->cus = prop.multiProcessorCount;
    b->tx = R == 4 && p->tile_x != 64 ? 32 : 64;  // as erm_create: 32 x 8 by default at radius 4
    b->ty = b->tx == 64 ? (p->tile_y == 4 ? 4 : 8) : (p->tile_y == 16 ? 16 : 8);
    if (p->zchunk > 0) b->kchunk = p->zchunk;
    *out = b;
    return 0;
}

int erm_batch_destroy(erm_batch* b) {
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_shots(b);
    free_tables(b);
    if (b->d_it) (void)hipFree(b->d_it);
    if (b->d_stack) (void)hipFree(b->d_stack);
    for (auto& e : b->ev) if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int erm_batch_load_model(erm_batch* b, int which, const float* vp, const float* vs, const float* rho) {
    if (!b || which < 0 || which > 1 || !vp || !vs || !rho) return b ? b->fail(ERM_E_PARAM, "bad arguments") : ERM_E_PARAM;
    const size_t n = (size_t)b->p.nx * b->p.ny * b->p.nz;
    b->model[which][0].assign(vp, vp + n);
    b->model[which][1].assign(vs, vs + n);
    b->model[which][2].assign(rho, rho + n);
    b->have_model[which] = true;
    for (size_t s = 0; s < b->shots.size(); ++s) {
        int rc = load_window(b, (int)s, which);
        if (rc) return rc;
        if (which == 0) b->have_rec[s] = 0;  // records of the old true model are stale
    }
    b->tables_dirty = true;
    return 0;
}

int erm_batch_set_shots(erm_batch* b, const erm_shot_window* shots, int nshots) {
    if (!b || !shots || nshots <= 0) return b ? b->fail(ERM_E_PARAM, "bad arguments") : ERM_E_PARAM;
    const erm_params& P = b->p;
    for (int s = 0; s < nshots; ++s) {
        const erm_shot_window& w = shots[s];
        if (w.wx <= 0 || w.wy <= 0 || w.x0 < 0 || w.y0 < 0 || w.x0 + w.wx > P.nx || w.y0 + w.wy > P.ny)
            return b->fail(ERM_E_PARAM, "shot %d: window outside the model", s);
        if (w.sx < w.x0 || w.sx >= w.x0 + w.wx || w.sy < w.y0 || w.sy >= w.y0 + w.wy || w.sz < 0 || w.sz >= P.nz)
            return b->fail(ERM_E_PARAM, "shot %d: source outside its window", s);
    }
    BHC(b, hipSetDevice(b->device));
    free_shots(b);
    b->win.assign(shots, shots + nshots);
    for (int s = 0; s < nshots; ++s) {
        const erm_shot_window& w = shots[s];
        erm_params q = P;
        q.nx = w.wx;
        q.ny = w.wy;
        q.shot_start_x = w.sx - w.x0; q.shot_interval_x = 0;
        q.shot_start_y = w.sy - w.y0; q.shot_interval_y = 0;
        q.shot_start_z = w.sz; q.shot_interval_z = 0;
        const int rix = std::max(1, P.receiver_interval_x), riy = std::max(1, P.receiver_interval_y);
        q.receiver_start_x = 0; q.receiver_interval_x = rix; q.receiver_num_x = (w.wx + rix - 1) / rix;
        q.receiver_start_y = 0; q.receiver_interval_y = riy; q.receiver_num_y = (w.wy + riy - 1) / riy;
        q.tile_x = b->tx;
        q.tile_y = b->ty;
        q.zchunk = b->kchunk;
        q.xcd_map = P.xcd_map == 2 ? 2 : 0;  // band order of the window's own one-shot tile lists
        q.poynting = 0;                      // the batched steps keep no Poynting vectors
        erm_ctx* c = nullptr;
        int rc = erm_create(&q, b->device, &c);
        if (rc) { free_shots(b); return b->fail(rc, "shot %d: %s", s, erm_last_error(nullptr)); }
        b->shots.push_back(c);
        b->have_rec.push_back(0);
        if ((rc = erm_set_shot(c, 0, 0, 0))) return b->fail(rc, "shot %d: %s", s, erm_last_error(c));
        for (int which = 0; which < 2; ++which)
            if (b->have_model[which] && (rc = load_window(b, s, which))) return rc;
    }
    b->tables_dirty = true;
    return 0;
}

int erm_batch_set_launch_shape(erm_batch* b, int tile_x, int tile_y, int zchunk, int grid_blocks) {
    if (!b) return ERM_E_PARAM;
    const bool ok = (tile_x == 64 && (tile_y == 4 || tile_y == 8)) || (tile_x == 32 && b->R == 4 && (tile_y == 8 || tile_y == 16));
    if (!ok || zchunk < 1 || grid_blocks < 0)
        return b->fail(ERM_E_PARAM, "tile 64 x {4, 8} or (radius 4) 32 x {8, 16}, zchunk >= 1, grid_blocks >= 0");
    b->tx = tile_x;
    b->ty = tile_y;
    b->kchunk = zchunk;
    b->grid = grid_blocks;
    b->tables_dirty = true;
    return 0;
}

int erm_batch_set_launch(erm_batch* b, int tile_y, int zchunk, int grid_blocks) {
    return erm_batch_set_launch_shape(b, 64, tile_y, zchunk, grid_blocks);
}

int erm_batch_use_graph(erm_batch* b, int on) {
    if (!b) return ERM_E_PARAM;
    b->use_graph = on != 0;
    return 0;
}

int erm_batch_autotune_shape(erm_batch* b, int steps, int* tile_x, int* tile_y, int* zchunk, int* grid_blocks,
                             double* ms_per_step) {
    if (!b || steps <= 0) return b ? b->fail(ERM_E_PARAM, "steps must be positive") : ERM_E_PARAM;
    if (b->shots.empty()) return b->fail(ERM_E_STATE, "no shots");
    if (!b->have_model[1]) return b->fail(ERM_E_STATE, "smoothed model not loaded");
    BHC(b, hipSetDevice(b->device));
    const int tx0 = b->tx, ty0 = b->ty, kc0 = b->kchunk, gr0 = b->grid;
    double best = 1e300;
    int bx = tx0, bt = ty0, bk = kc0, bg = gr0;
    std::vector<std::pair<int, int>> shapes = {{64, 4}, {64, 8}};
    if (b->R == 4) { shapes.push_back({32, 8}); shapes.push_back({32, 16}); }
    const int kcs[4] = {4, 6, 8, 16};
    for (const auto& sh : shapes)
        for (int kc : kcs) {
            const int tx = sh.first, ty = sh.second;
            // resident workgroups: 16 waves per CU at 4 waves/SIMD
            const int per_cu = std::max(1, 16 * 64 / (tx * ty));
            const int grids[3] = {0, b->cus * per_cu, 2 * b->cus * per_cu};
            bool fits = true;
            for (erm_ctx* c : b->shots) fits = fits && sweep_fits(c, std::max(1, std::min(kc, c->nnz - 2 * c->R)));
            if (!fits) continue;
            for (int gr : grids) {
                b->tx = tx; b->ty = ty; b->kchunk = kc; b->grid = gr;
                b->tables_dirty = true;
                int rc;
                if ((rc = zero_shots(b, true, true, false))) return rc;
                float ms = 0;
                if ((rc = run_steps(b, ERM_PASS_SOURCE, 0, 1, false, &ms))) return rc;       // warm-up
                if ((rc = run_steps(b, ERM_PASS_SOURCE, 1, steps, false, &ms))) return rc;
                const double per = ms / steps;
                if (per < best) { best = per; bx = tx; bt = ty; bk = kc; bg = gr; }
            }
        }
    b->tx = bx; b->ty = bt; b->kchunk = bk; b->grid = bg;
    b->tables_dirty = true;
    int rc;
    if (best == 1e300) {  // no candidate fits the 32-bit span: time the current shape
        float ms = 0;
        if ((rc = zero_shots(b, true, true, false)) || (rc = run_steps(b, ERM_PASS_SOURCE, 0, 1, false, &ms)) ||
            (rc = run_steps(b, ERM_PASS_SOURCE, 1, steps, false, &ms)))
            return rc;
        best = ms / steps;
    }
    rc = zero_shots(b, true, true, false);
    if (rc) return rc;
    if (tile_x) *tile_x = bx;
    if (tile_y) *tile_y = bt;
    if (zchunk) *zchunk = bk;
    if (grid_blocks) *grid_blocks = bg;
    if (ms_per_step) *ms_per_step = best;
    return 0;
}

int erm_batch_autotune(erm_batch* b, int steps, int* tile_y, int* zchunk, int* grid_blocks, double* ms_per_step) {
    return erm_batch_autotune_shape(b, steps, nullptr, tile_y, zchunk, grid_blocks, ms_per_step);
}

int erm_batch_set_record(erm_batch* b, int shot, const float* rec) {
    if (!b || shot < 0 || shot >= (int)b->shots.size() || !rec) return b ? b->fail(ERM_E_PARAM, "bad arguments") : ERM_E_PARAM;
    int rc = erm_set_record(b->shots[shot], rec);
    if (rc) return b->fail(rc, "shot %d: %s", shot, erm_last_error(b->shots[shot]));
    b->have_rec[shot] = 1;
    return 0;
}

int erm_batch_run(erm_batch* b, float* stp, float* spp, float* sps) {
    if (!b) return ERM_E_PARAM;
    if (b->shots.empty()) return b->fail(ERM_E_STATE, "no shots");
    if (!b->have_model[1]) return b->fail(ERM_E_STATE, "smoothed model not loaded");
    BHC(b, hipSetDevice(b->device));
    int rc;
    bool need_model = false;
    for (char h : b->have_rec) need_model |= !h;
    b->ms[0] = b->ms[1] = b->ms[2] = 0;
    b->steps[0] = b->steps[1] = b->steps[2] = 0;
    float ms = 0;
    if (need_model) {
        // the model pass runs for all shots at once: all or none must hold records
        for (char h : b->have_rec)
            if (h) return b->fail(ERM_E_STATE, "either every shot has a record or none");
        if (!b->have_model[0]) return b->fail(ERM_E_STATE, "true model not loaded and no records set");
        if ((rc = zero_shots(b, true, false, false))) return rc;
        if ((rc = run_steps(b, ERM_PASS_MODEL, 0, b->p.lt, b->use_graph, &ms))) return rc;
        b->ms[0] = ms;
        b->steps[0] = b->p.lt;
        if (hipStreamSynchronize(b->stream) != hipSuccess) return b->fail(ERM_E_DEVICE, "model pass failed");
        if (b->progress) b->progress(ERM_PASS_MODEL, ms, b->progress_user);
        for (size_t s = 0; s < b->shots.size(); ++s) {
            if ((rc = mute_direct(b->shots[s])))  // as erm_forward_real: the window's own direct wave
                return b->fail(rc, "shot %zu: %s", s, b->shots[s]->err.c_str());
            b->shots[s]->have_record = true;
            b->have_rec[s] = 1;  // kept for later runs until the true model or the shots change
        }
    }
    if ((rc = zero_shots(b, true, true, false))) return rc;
    if ((rc = run_steps(b, ERM_PASS_SOURCE, 0, b->p.lt_forward, b->use_graph, &ms))) return rc;
    b->ms[1] = ms;
    b->steps[1] = b->p.lt_forward;
    if (b->progress) b->progress(ERM_PASS_SOURCE, ms, b->progress_user);
    if ((rc = zero_shots(b, true, false, true))) return rc;
    if ((rc = run_steps(b, ERM_PASS_RECEIVER, b->p.lt - 1, b->p.lt, b->use_graph, &ms))) return rc;
    b->ms[2] = ms;
    b->steps[2] = b->p.lt;
    if (b->progress) b->progress(ERM_PASS_RECEIVER, ms, b->progress_user);
    // stack the window images into the full model
    const erm_params& P = b->p;
    const size_t nfull = (size_t)P.nx * P.ny * P.nz;
    if (!b->d_stack) BHC(b, hipMalloc(&b->d_stack, 3 * nfull * sizeof(float)));
    hipLaunchKernelGGL(k_fill_f32, dim3(1024), dim3(256), 0, b->stream, b->d_stack, (long)(3 * nfull), 0.f);
    for (size_t s = 0; s < b->shots.size(); ++s) {
        erm_ctx* c = b->shots[s];
        const erm_shot_window& w = b->win[s];
        const long n = (long)w.wx * w.wy * P.nz;
        for (int q = 0; q < 3; ++q)
            hipLaunchKernelGGL(k_stack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, c->img[q], w.wx, w.wy,
                               P.nz, P.bl, P.bf, P.bu, c->px, c->pxy, b->d_stack + q * nfull, P.nx, P.ny, w.x0, w.y0);
    }
    BHC(b, hipGetLastError());
    float* outs[3] = {stp, spp, sps};
    std::vector<float> tmp(nfull);
    for (int q = 0; q < 3; ++q) {
        if (!outs[q]) continue;
        BHC(b, hipMemcpyAsync(tmp.data(), b->d_stack + q * nfull, nfull * sizeof(float), hipMemcpyDeviceToHost, b->stream));
        BHC(b, hipStreamSynchronize(b->stream));
        for (int k = 0; k < P.nz; ++k)
            for (int j = 0; j < P.ny; ++j)
                for (int i = 0; i < P.nx; ++i)
                    outs[q][((size_t)i * P.ny + j) * P.nz + k] = tmp[((size_t)k * P.ny + j) * P.nx + i];
    }
    BHC(b, hipStreamSynchronize(b->stream));
    return 0;
}

int erm_batch_set_progress(erm_batch* b, erm_batch_progress_fn fn, void* user) {
    if (!b) return ERM_E_PARAM;
    b->progress = fn;
    b->progress_user = user;
    return 0;
}

int erm_batch_shot_context(erm_batch* b, int shot, erm_ctx** ctx) {
    if (!b || !ctx || shot < 0 || shot >= (int)b->shots.size()) return b ? b->fail(ERM_E_PARAM, "bad shot") : ERM_E_PARAM;
    *ctx = b->shots[shot];
    return 0;
}

int erm_batch_timing(const erm_batch* b, double* m0, double* m1, double* m2, double* cells, int* s0, int* s1, int* s2) {
    if (!b) return ERM_E_PARAM;
    if (m0) *m0 = b->ms[0];
    if (m1) *m1 = b->ms[1];
    if (m2) *m2 = b->ms[2];
    if (cells) *cells = b->cells_per_step;
    if (s0) *s0 = b->steps[0];
    if (s1) *s1 = b->steps[1];
    if (s2) *s2 = b->steps[2];
    return 0;
}

}  // extern "C"
