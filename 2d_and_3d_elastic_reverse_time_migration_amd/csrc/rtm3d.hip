// rtm3d.hip - the 3D elastic RTM context behind the C ABI of include/erm.h.
//
// Replaces the reference's GPUdevice struct and allocate/zero helpers
// (3D_elastic_RTM/3D_elastic_modeling_typedef_struct:1-293), the three host
// time loops (3D_forward_or_back_together.cu:9-155,161-290,296-453) and the
// result writers (3D_output_file.cu:155-250,309-428) with one device-resident
// context per GPU.  Differences by design:
//  - the whole padded grid lives on one GPU (no z-slab split, no halo
//    exchange: the reference's exchange_device_nz_kernel_* are not needed);
//  - the receiver record stays in HBM for the whole shot instead of one
//    time slice crossing PCIe per step (:140,318);
//  - two fused kernels per step, no device-wide synchronisation inside a pass.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "common.hpp"
#include "ctx3d.hpp"
#include "kernels3d.hpp"

namespace erm {

static thread_local std::string g_err;

std::string vformat(const char* fmt, va_list ap) {
    char buf[1024];
    vsnprintf(buf, sizeof(buf), fmt, ap);
    return std::string(buf);
}
std::string format(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::string s = vformat(fmt, ap);
    va_end(ap);
    return s;
}
void set_global_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    g_err = vformat(fmt, ap);
    va_end(ap);
}
const char* global_error() { return g_err.c_str(); }

// Reference radius-6 optimized coefficients (3D_elastic_RTM/zzzzz:205-210)
// and radius-4 staggered Taylor coefficients.
static const double kCoeR6[7] = {0.0, 1.2213363647, -0.096931457519, 0.017447662353,
                                 -0.0029672895159, 0.00035900539822, -0.00002184781161};
static const double kCoeR4[5] = {0.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0};

}  // namespace erm

using namespace erm;

namespace {

// 1-cos face profile of coe_attenuation_3d_new2 (3D_zzzzz:284-313), pai = 3.1415926.
void face_profile(int b, double cA, bool rising, std::vector<double>& d) {
    const double pai = 3.1415926;
    d.assign(b, 0.0);
    for (int i = 0; i < b; ++i) {
        double s = rising ? (double)i / (b - 1) : (double)(b - i - 1) / (b - 1);
        d[i] = cA * (1.0 - std::cos(pai / 2 * s));
    }
}

// make_ricker_new (3D_elastic_RTM/zzzzz:136-157), dt in seconds, cwp PI.
std::vector<float> ricker(float freq, float dt_s) {
    const double PI = 3.141592653589793;
    if (freq == 0.0f) freq = 30.0f;
    if (dt_s == 0.0f) dt_s = 0.004f;
    float B = (float)(std::sqrt(6.0) / (PI * freq));
    int N = (int)std::ceil(1.35 * (double)B / (double)dt_s);
    std::vector<float> w(2 * N + 1);
    w[N] = 1.0f;
    for (int i = 1; i <= N; ++i) {
        float t = dt_s * (float)i;
        float u = (float)(2.0 * std::sqrt(6.0)) * t / B;
        w[N + i] = w[N - i] = 0.5f * (2.0f - u * u) * std::exp(-u * u / 4.0f);
    }
    return w;
}


template <typename T>
int dmalloc(erm_ctx* c, T** ptr, size_t n) {
    if (n == 0) { *ptr = nullptr; return 0; }
    hipError_t e = hipMalloc((void**)ptr, n * sizeof(T));
    if (e != hipSuccess) return c->fail(ERM_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", n * sizeof(T), hipGetErrorString(e));
    return 0;
}

// a volume (ncell floats, the x = R column of every row 128-byte aligned
// when xoff > 0): the allocation holds xoff floats in front of the base
int vmalloc(erm_ctx* c, float** ptr, size_t n) {
    float* p = nullptr;
    if (int rc = dmalloc(c, &p, n + c->xoff)) return rc;
    *ptr = p + c->xoff;
    return 0;
}
void vfree(const erm_ctx* c, float* p) {
    if (p) (void)hipFree(p - c->xoff);
}

inline dim3 grid_of(const Geo& g) { return dim3(g.xcd_map ? 8 * g.per_xcd : g.ntiles); }

// Tile groups of the CPML region split for the current launch shape: group 0
// holds the tiles none of whose updated cells lies in a memory-variable slab
// (they run the plain kernel), group 7 all others (the all-axes C-PML kernel).
// Tiles with no updated cell are dropped.  The kernels also build for the
// single-axis groups 1..6 (bit 0 x, 1 y, 2 z), but launching the groups one
// after another measured slower (launch tails) than the two-group split.
// XCD-aware order of a tile list: workgroup b runs on XCD b % 8, each XCD with
// its own L2, so XCD x gets the x-th contiguous chunk of the (x-fastest) list
// -- its tiles' x, y and z neighbours, whose halo planes and rows it re-reads,
// then sit in the same L2 -- at positions x, x + 8, x + 16, ... (chunk sizes
// differ by at most one; the map is a bijection of [0, n)).
std::vector<int> xcd_order(const std::vector<int>& t) {
    const int n = (int)t.size(), s = n / 8, r = n % 8;
    std::vector<int> out(n);
    for (int b = 0; b < n; ++b) {
        const int x = b % 8, j = b / 8;
        out[b] = t[x * s + std::min(x, r) + j];
    }
    return out;
}

// XCD y-band order of a tile list (xcd_map 2): the list, sorted y-major
// (by, bz, bx), is cut into 8 chunks of equal length, so XCD x owns a
// contiguous band of rows of the grid; inside its chunk the tiles run
// (bz, by, bx), so the workgroups an XCD holds at once form a compact patch
// of one z layer whose x and y halos its own L2 serves.  Chunk x fills
// positions x, x + 8, ...; positions past a shorter chunk hold the sentinel
// `none` (a tile id that the kernels skip).
std::vector<int> xcd_band_order(const std::vector<int>& t, const Geo& g, int none) {
    const int n = (int)t.size();
    auto key = [&](int l, bool ymajor) {
        const long bx = l % g.ntx, by = (l / g.ntx) % g.nty, bz = l / (g.ntx * g.nty);
        return ymajor ? (by * g.ntiles + bz) * g.ntx + bx : (bz * g.nty + by) * g.ntx + bx;
    };
    std::vector<int> s = t;
    std::sort(s.begin(), s.end(), [&](int a, int b) { return key(a, true) < key(b, true); });
    const int per = (n + 7) / 8;
    std::vector<int> out((size_t)8 * per, none);
    for (int x = 0; x < 8; ++x) {
        const int b0 = std::min(n, x * per), b1 = std::min(n, b0 + per);
        std::sort(s.begin() + b0, s.begin() + b1, [&](int a, int b) { return key(a, false) < key(b, false); });
        for (int j = 0; j < b1 - b0; ++j) out[(size_t)8 * j + x] = s[b0 + j];
    }
    return out;
}

int build_split(erm_ctx* c) {
    if (c->split_built && c->split_tx == c->tx && c->split_ty == c->ty && c->split_kchunk == c->kchunk) return 0;
    Geo g = make_geo(c);
    const int R = c->R;
    std::vector<int> grp[8];
    std::vector<int> axes(g.ntiles, 0);  // per tile: the axes whose slabs it touches
    if (g.ntiles > kTileIdMask) return c->fail(ERM_E_SIZE, "too many tiles for the tile lists");
    auto inside = [](int lo_c, int hi_c, int n, int lo, int hi) {  // [lo_c, hi_c] avoids both slabs
        return lo_c > lo && hi_c < n - hi - 1;
    };
    for (int l = 0; l < g.ntiles; ++l) {
        const int bx = l % g.ntx, by = (l / g.ntx) % g.nty, bz = l / (g.ntx * g.nty);
        const int i0 = std::max(c->x0 + bx * c->tx, R), i1 = std::min(c->x0 + bx * c->tx + c->tx, c->nnx - R) - 1;
        const int j0 = std::max(by * c->ty, R), j1 = std::min(by * c->ty + c->ty, c->nny - R) - 1;
        const ZSpan zs = chunk_span(g, R, bz);
        const int k0 = zs.kb, k1 = zs.ke - 1;
        if (i0 > i1 || j0 > j1 || k0 > k1) continue;
        int m = (inside(i0, i1, c->nnx, c->slab_lo[0], c->slab_hi[0]) ? 0 : 1) |
                (inside(j0, j1, c->nny, c->slab_lo[1], c->slab_hi[1]) ? 0 : 2) |
                (inside(k0, k1, c->nnz, c->slab_lo[2], c->slab_hi[2]) ? 0 : 4);
        axes[l] = m;
        if (m != 0) m = 7;
        grp[m].push_back(l);
    }
    c->split_built = false;  // until every list below is in place
    for (int m = 0; m < 8; ++m) {
        if (c->tiles[m]) (void)hipFree(c->tiles[m]);
        c->tiles[m] = nullptr;
        if (!grp[m].empty() && c->xcd_map == 1) grp[m] = xcd_order(grp[m]);
        if (!grp[m].empty() && c->xcd_map == 2) grp[m] = xcd_band_order(grp[m], g, g.ntiles);
        // C-PML tiles carry their axis mask (bits 28..30) so the kernel skips the
        // memory variables of the axes whose slabs the tile does not touch
        for (int& t : grp[m]) if (t < g.ntiles) t |= axes[t] << 28;
        c->ntl[m] = (int)grp[m].size();
        if (grp[m].empty()) continue;
        if (dmalloc(c, &c->tiles[m], grp[m].size())) return ERM_E_NOMEM;
        HIPC(c, hipMemcpy(c->tiles[m], grp[m].data(), grp[m].size() * sizeof(int), hipMemcpyHostToDevice));
    }
    c->split_tx = c->tx;
    c->split_ty = c->ty;
    c->split_kchunk = c->kchunk;
    c->split_built = true;
    return 0;
}

template <int R, int TX, int TY, int AX>
void stress_one(erm_ctx* c, int which, int src_k, int src_col, int it, const Geo& g, dim3 grd, const int* list,
                hipStream_t st = nullptr) {
    hipLaunchKernelGGL((k_stress<R, TX, TY, AX>), grd, dim3(TX, TY), 0, st ? st : c->stream, make_wave(c),
                       c->M[which], c->MU[which], g, c->coef, src_k, src_col, c->wavelet, c->wl, it, c->pml, list);
}
template <int R, int TX, int TY, int MODE, int AX>
void velocity_one(erm_ctx* c, int which, int it, const Geo& g, dim3 grd, const int* list, hipStream_t st = nullptr) {
    hipLaunchKernelGGL((k_velocity<R, TX, TY, MODE, AX>), grd, dim3(TX, TY), 0, st ? st : c->stream, make_wave(c),
                       c->BU[which], g, c->coef, make_excite(c), it, c->pml, list);
}
// region launches of one step: the C-PML group on the second stream (when
// region_streams), concurrent with the plain group on the context's stream;
// fork before, join after, so the step is ordered like one launch
hipStream_t region_fork(erm_ctx* c) {
    if (!c->stream2) return c->stream;
    (void)hipEventRecord(c->fork_ev, c->stream);
    (void)hipStreamWaitEvent(c->stream2, c->fork_ev, 0);
    return c->stream2;
}
void region_join(erm_ctx* c) {
    if (!c->stream2) return;
    (void)hipEventRecord(c->join_ev, c->stream2);
    (void)hipStreamWaitEvent(c->stream, c->join_ev, 0);
}

template <int R, int TX, int TY>
void stress_launch(erm_ctx* c, int which, long src, int it) {
    const int src_k = src < 0 ? -1 : (int)(src / c->pxy), src_col = src < 0 ? -1 : (int)(src % c->pxy);
    Geo g = make_geo(c);
    if (c->cpml && c->region_split && build_split(c) == 0) {
        g.xcd_map = 0;
        const hipStream_t s2 = region_fork(c);
        if (c->ntl[7]) stress_one<R, TX, TY, 7>(c, which, src_k, src_col, it, g, dim3(c->ntl[7]), c->tiles[7], s2);
        if (c->ntl[0]) stress_one<R, TX, TY, 0>(c, which, src_k, src_col, it, g, dim3(c->ntl[0]), c->tiles[0]);
        region_join(c);
    } else if (c->cpml) {
        stress_one<R, TX, TY, 7>(c, which, src_k, src_col, it, g, grid_of(g), nullptr);
    } else {
        stress_one<R, TX, TY, 0>(c, which, src_k, src_col, it, g, grid_of(g), nullptr);
    }
}

template <int R, int TX, int TY, int MODE>
void velocity_launch(erm_ctx* c, int which, int it) {
    Geo g = make_geo(c);
    if (c->cpml && c->region_split && build_split(c) == 0) {
        g.xcd_map = 0;
        const hipStream_t s2 = region_fork(c);
        if (c->ntl[7]) velocity_one<R, TX, TY, MODE, 7>(c, which, it, g, dim3(c->ntl[7]), c->tiles[7], s2);
        if (c->ntl[0]) velocity_one<R, TX, TY, MODE, 0>(c, which, it, g, dim3(c->ntl[0]), c->tiles[0]);
        region_join(c);
    } else if (c->cpml) {
        velocity_one<R, TX, TY, MODE, 7>(c, which, it, g, grid_of(g), nullptr);
    } else {
        velocity_one<R, TX, TY, MODE, 0>(c, which, it, g, grid_of(g), nullptr);
    }
}

// ---------------------------------------------------------------------------
// Source-wavefield reconstruction.  Inside the model region shrunk by one
// cell (the first model cells of some faces still see a non-zero C-PML
// profile at their staggered positions) the update has no damping and no
// memory variables, so a step is inverted by the same kernels with dt
// negated: v^n = v^{n+1} - dt b D sigma^{n+1/2}, then
// sigma^{n-1/2} = sigma^{n+1/2} - dt C D v^n - wavelet[n] (the source enters
// before the update).  The stencils of those cells reach R cells out, into
// the face slabs whose forward values (sigma + tp after the stress update;
// v + vP after the velocity update) were saved every step; they are restored
// before each reversed half-step, so every model-region cell holds the
// forward value when the images are formed.  Cells further out are not
// needed (they are updated without absorption and stay bounded).
// Strip layout: slot s holds 13 fields x n_strip; slot it+1 is forward step
// it (7 stress-side fields, then 6 velocity-side fields); slot 0 is zero.
// ---------------------------------------------------------------------------
constexpr int kStripFields = 13;
inline Wave wave_of(float* const* f) {
    Wave w;
    w.vx = f[0]; w.vy = f[1]; w.vz = f[2];
    w.txx = f[3]; w.tyy = f[4]; w.tzz = f[5]; w.txy = f[6]; w.txz = f[7]; w.tyz = f[8];
    w.tp = f[9]; w.vxp = f[10]; w.vyp = f[11]; w.vzp = f[12];
    return w;
}
inline Geo reversed_geo(const erm_ctx* c) {
    Geo g = make_geo(c);
    g.cx = -g.cx; g.cy = -g.cy; g.cz = -g.cz;
    return g;
}
inline StripFields stress_side(float* const* f) { return StripFields{{f[3], f[4], f[5], f[6], f[7], f[8], f[9]}}; }
inline StripFields velocity_side(float* const* f) { return StripFields{{f[0], f[1], f[2], f[10], f[11], f[12], nullptr}}; }
// restore (or save) the stress-side slab values of level it+1/2 = forward step it
void strip_stress(erm_ctx* c, float* const* f, int it, int restore) {
    const int n = c->n_strip, rb = 256, rg = (n + rb - 1) / rb;
    hipLaunchKernelGGL(k_strip, dim3(rg), dim3(rb), 0, c->stream, stress_side(f), 7, (const int*)c->strip_idx, n,
                       c->strips + (long)(it + 1) * kStripFields * n, restore);
}
// restore (or save) the velocity-side slab values of level lev (= forward step lev-1; lev 0 is zero)
void strip_velocity(erm_ctx* c, float* const* f, int lev, int restore) {
    const int n = c->n_strip, rb = 256, rg = (n + rb - 1) / rb;
    hipLaunchKernelGGL(k_strip, dim3(rg), dim3(rb), 0, c->stream, velocity_side(f), 6, (const int*)c->strip_idx, n,
                       c->strips + (long)lev * kStripFields * n + 7L * n, restore);
}
// reverse step it after the stress-side slabs of level it+1/2 are in place:
// state {sigma^{it+1/2}, v^{it+1}} -> {sigma^{it-1/2}, v^{it}}
template <int R, int TX, int TY>
void reverse_rest_rt(erm_ctx* c, int it) {
    const Geo g = reversed_geo(c);
    const Wave w = wave_of(c->fr);
    hipLaunchKernelGGL((k_velocity<R, TX, TY, MODE_PLAIN, 0>), grid_of(g), dim3(TX, TY), 0, c->stream, w, c->BU[1], g,
                       c->coef, make_excite(c), it, c->pml, nullptr);
    strip_velocity(c, c->fr, it, 1);
    const int src_k = c->src < 0 ? -1 : (int)(c->src / c->pxy), src_col = c->src < 0 ? -1 : (int)(c->src % c->pxy);
    hipLaunchKernelGGL((k_stress<R, TX, TY, 0>), grid_of(g), dim3(TX, TY), 0, c->stream, w, c->M[1], c->MU[1], g,
                       c->coef, src_k, src_col, (const float*)c->wavelet_neg, c->wl, it, c->pml, nullptr);
}
void reverse_rest(erm_ctx* c, int it) {
    if (c->R == 4) {
        if (c->tx == 32) { if (c->ty == 8) reverse_rest_rt<4, 32, 8>(c, it); else reverse_rest_rt<4, 32, 16>(c, it); }
        else if (c->ty == 4) reverse_rest_rt<4, 64, 4>(c, it);
        else reverse_rest_rt<4, 64, 8>(c, it);
    } else {
        if (c->ty == 4) reverse_rest_rt<6, 64, 4>(c, it); else reverse_rest_rt<6, 64, 8>(c, it);
    }
}
void reverse_step(erm_ctx* c, int it) {
    strip_stress(c, c->fr, it, 1);
    reverse_rest(c, it);
}

// C-PML coefficients (Roden & Gedney 2000; Komatitsch & Martin 2007, kappa = 1):
// b = exp(-(d+alpha) dt), a = d (b-1)/(d+alpha), d = d0 r^2,
// d0 = -3 vmax ln(Rc) / (2 L), alpha = pi f0 (1 - r), r = depth into the layer / L,
// L = layer width - R (the outer R cells are never updated),
// at integer positions and at the staggered ones (x-1/2, y-1/2, z+1/2).
// Also counts, per kernel, the memory-variable bytes the launch moves.
int build_cpml(erm_ctx* c) {
    const erm_params& p = c->p;
    const double vmax = p.cpml_vmax > 0 ? p.cpml_vmax : c->vmax_loaded;
    const double Rc = p.cpml_r > 0 ? p.cpml_r : 1e-3;
    const double dts = p.dt / 1000.0;
    const int n3[3] = {c->nnx, c->nny, c->nnz};
    const int lo[3] = {p.bl, p.bf, p.bu}, hi[3] = {p.br, p.bb, p.bd};
    const double hh[3] = {p.dx, p.dy, p.dz};
    const double shift[3] = {-0.5, -0.5, +0.5};
    // per axis, per index: (a_full, b_full, a_half, b_half); a = 0, b = 1 outside the layers
    std::vector<float> host(4 * (size_t)(n3[0] + n3[1] + n3[2]));
    auto at = [&](int a, int f, int ab, int i) -> float& {
        return host[coef_base(c->nnx, c->nny, a) + 4 * (size_t)i + 2 * f + ab];
    };
    for (int a = 0; a < 3; ++a)
        for (int f = 0; f < 2; ++f) {
            for (int i = 0; i < n3[a]; ++i) { at(a, f, 0, i) = 0.f; at(a, f, 1, i) = 1.f; }
            for (int i = 0; i < n3[a]; ++i) {
                // the layer spans the updated cells only (the outer R never update)
                double x = i + (f ? shift[a] : 0.0), s = 0, L = 0;
                if (lo[a] > 0 && x < lo[a]) { s = lo[a] - x; L = std::max(lo[a] - c->R, 1); }
                else if (hi[a] > 0 && x > n3[a] - hi[a] - 1) { s = x - (n3[a] - hi[a] - 1); L = std::max(hi[a] - c->R, 1); }
                if (L <= 0) continue;
                double r = s / L, d0 = -3.0 * vmax * std::log(Rc) / (2.0 * L * hh[a]);
                double d = d0 * r * r, al = M_PI * p.freq * (1.0 - std::min(r, 1.0));
                double bb = std::exp(-(d + al) * dts);
                at(a, f, 1, i) = (float)bb;
                at(a, f, 0, i) = (float)(d > 0 ? d * (bb - 1.0) / (d + al) : 0.0);
            }
        }
    float* dcoef = c->psi_buf + psi_total(c->pml);
    HIPC(c, hipMemcpyAsync(dcoef, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    // memory-variable traffic: 8 B (read+write) per psi at every updated cell whose profile is non-zero
    const int R = c->R;
    double cnt[3][2] = {{0, 0}, {0, 0}, {0, 0}};  // updated positions with a != 0 along each axis
    for (int a = 0; a < 3; ++a)
        for (int f = 0; f < 2; ++f)
            for (int i = R; i < n3[a] - R; ++i)
                if (at(a, f, 0, i) != 0.f) cnt[a][f] += 1;
    const double span[3] = {(double)(c->nnx - 2 * R), (double)(c->nny - 2 * R), (double)(c->nnz - 2 * R)};
    // stress psi per axis: 1 full + 2 half; velocity psi per axis: 2 full + 2 half
    double st = 0, ve = 0;
    for (int a = 0; a < 3; ++a) {
        st += 8.0 * (cnt[a][0] + 2 * cnt[a][1]) / span[a];
        ve += 8.0 * (2 * cnt[a][0] + 2 * cnt[a][1]) / span[a];
    }
    c->pml_extra_stress = st;
    c->pml_extra_velocity = ve;
    return 0;
}

// launch shapes: 64 x {4, 8} at both radii, 32 x {8, 16} at radius 4
void do_stress(erm_ctx* c, int which, long src, int it) {
    if (c->R == 4) {
        if (c->tx == 32) {
            if (c->ty == 8) stress_launch<4, 32, 8>(c, which, src, it);
            else stress_launch<4, 32, 16>(c, which, src, it);
        } else if (c->ty == 4) stress_launch<4, 64, 4>(c, which, src, it);
        else stress_launch<4, 64, 8>(c, which, src, it);
    } else {
        if (c->ty == 4) stress_launch<6, 64, 4>(c, which, src, it);
        else stress_launch<6, 64, 8>(c, which, src, it);
    }
}

template <int MODE>
void do_velocity_m(erm_ctx* c, int which, int it) {
    if (c->R == 4) {
        if (c->tx == 32) {
            if (c->ty == 8) velocity_launch<4, 32, 8, MODE>(c, which, it);
            else velocity_launch<4, 32, 16, MODE>(c, which, it);
        } else if (c->ty == 4) velocity_launch<4, 64, 4, MODE>(c, which, it);
        else velocity_launch<4, 64, 8, MODE>(c, which, it);
    } else {
        if (c->ty == 4) velocity_launch<6, 64, 4, MODE>(c, which, it);
        else velocity_launch<6, 64, 8, MODE>(c, which, it);
    }
}

void do_velocity(erm_ctx* c, int which, int mode, int it) {
    if (mode == MODE_EXCITE) do_velocity_m<MODE_EXCITE>(c, which, it);
    else if (mode == MODE_IMAGE) do_velocity_m<MODE_IMAGE>(c, which, it);
    else do_velocity_m<MODE_PLAIN>(c, which, it);
}

int fill(erm_ctx* c, float* p, size_t n, float v) {
    if (!p) return 0;
    hipLaunchKernelGGL(k_fill_f32, dim3(1024), dim3(256), 0, c->stream, p, (long)n, v);
    HIPC(c, hipGetLastError());
    return 0;
}
int filli(erm_ctx* c, int* p, size_t n, int v) {
    if (!p) return 0;
    hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, c->stream, p, (long)n, v);
    HIPC(c, hipGetLastError());
    return 0;
}

int check_grid(const erm_params* p, int R, std::string* err) {
    if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0) { *err = "nx, ny, nz must be positive"; return ERM_E_PARAM; }
    if (p->dx <= 0 || p->dy <= 0 || p->dz <= 0 || p->dt <= 0) { *err = "dx, dy, dz, dt must be positive"; return ERM_E_PARAM; }
    if (p->lt <= 0) { *err = "lt must be positive"; return ERM_E_PARAM; }
    int b[6] = {p->bu, p->bd, p->bl, p->br, p->bf, p->bb};
    for (int v : b)
        if (v < 0 || v == 1) { *err = "boundary widths must be 0 or >= 2"; return ERM_E_PARAM; }
    long nnx = p->nx + p->bl + p->br, nny = p->ny + p->bf + p->bb, nnz = p->nz + p->bu + p->bd;
    if (nnx < 2 * R + 1 || nny < 2 * R + 1 || nnz < 2 * R + 1) { *err = "padded grid smaller than the stencil"; return ERM_E_SIZE; }
    if (nnx * nny * nnz > (1L << 34)) { *err = "grid too large"; return ERM_E_SIZE; }
    if (p->receiver_num_x < 0 || p->receiver_num_y < 0) { *err = "receiver counts must be >= 0"; return ERM_E_PARAM; }
    if (p->boundary != ERM_BOUNDARY_DAMP && p->boundary != ERM_BOUNDARY_CPML) { *err = "boundary must be 0 (damping) or 1 (CPML)"; return ERM_E_PARAM; }
    if (p->join_wavefield != 0 && p->wavefield_interval <= 0) { *err = "join_wavefield needs wavefield_interval > 0"; return ERM_E_PARAM; }
    // reference keys whose code is not part of this library: refused when set
    // to anything but their neutral value instead of being ignored
    struct { const char* key; bool set; const char* what; } off[] = {
        {"laplace", p->laplace != 0, "Laplacian image filter (cpu_laplace, zzzzz:1110-1147, host post-processing)"},
        {"laplace_compensate", p->laplace_compensate != 0, "Laplacian compensation (no code in the reference)"},
        {"precon", p->precon != 0, "image preconditioning / precon_z1-z2 depth window (kernel bodies absent in the reference)"},
        {"vsp_precon", p->vsp_precon != 0, "VSP preconditioning (no code in the reference)"},
        {"velocity_multiple", p->velocity_multiple != 0.f && p->velocity_multiple != 1.f,
         "velocity scaling for model smoothing (SU smooth3d, outside the library: supply smoothed models)"},
    };
    for (const auto& o : off)
        if (o.set) { *err = std::string(o.key) + " is not supported: " + o.what; return ERM_E_PARAM; }
    return 0;
}


void free_all(erm_ctx* c) {
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    for (auto& p : c->f) vfree(c, p);
    for (int w = 0; w < 2; ++w) { vfree(c, c->M[w]); vfree(c, c->MU[w]); vfree(c, c->BU[w]); }
    for (auto& p : c->hxyz) F(p);
    F(c->wavelet); vfree(c, c->exbuf);
    F(c->rec); F(c->rec_idx);
    F(c->psi_buf);
    for (auto& t : c->tiles) F(t);
    for (auto& q : c->fr) vfree(c, q);
    F(c->strip_idx); F(c->strips); F(c->wavelet_neg); F(c->hist);
    for (auto& q : c->ximg) vfree(c, q);
    for (auto& q : c->poyn) vfree(c, q);
    vfree(c, c->decbuf);
    for (auto& e : c->pass_ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : c->kev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
}

// divergence (q = 0) or curl component q of the current velocity into decbuf
int decompose(erm_ctx* c, int q) {
    if (!c->decbuf && vmalloc(c, &c->decbuf, c->ncell)) return ERM_E_NOMEM;
    const long tot = (long)c->nnx * c->nny * c->nnz;
    const dim3 grd((unsigned)((tot + 255) / 256));
    const float ix = (float)(1.0 / c->p.dx), iy = (float)(1.0 / c->p.dy), iz = (float)(1.0 / c->p.dz);
    if (c->R == 4)
        hipLaunchKernelGGL(k_decom<4>, grd, dim3(256), 0, c->stream, c->f[0], c->f[1], c->f[2], c->decbuf, q, c->nnx,
                           c->nny, c->nnz, c->px, c->pxy, c->coef, ix, iy, iz);
    else
        hipLaunchKernelGGL(k_decom<6>, grd, dim3(256), 0, c->stream, c->f[0], c->f[1], c->f[2], c->decbuf, q, c->nnx,
                           c->nny, c->nnz, c->px, c->pxy, c->coef, ix, iy, iz);
    HIPC(c, hipGetLastError());
    return 0;
}

// download a padded-grid volume into a dense nnx*nny*nnz host array
int download(erm_ctx* c, const float* d, float* h) {
    HIPC(c, hipMemcpy2DAsync(h, (size_t)c->nnx * sizeof(float), d, (size_t)c->px * sizeof(float),
                             (size_t)c->nnx * sizeof(float), (size_t)c->nny * c->nnz, hipMemcpyDeviceToHost,
                             c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}
int upload(erm_ctx* c, float* d, const float* h) {
    HIPC(c, hipMemcpy2DAsync(d, (size_t)c->px * sizeof(float), h, (size_t)c->nnx * sizeof(float),
                             (size_t)c->nnx * sizeof(float), (size_t)c->nny * c->nnz, hipMemcpyHostToDevice,
                             c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}

// Moduli of the padded grid from the model region in disk order [nx][ny][nz]
// (z fastest), edge-replicated into the absorbing layers (add_pml_layers_v_h,
// 3D_zzzzz:141-259): one thread per padded cell, x fastest.  A missing vs is
// vp/sqrt(3) (vp_set_vs, zzzzz:1261-1270), a missing rho 230 vp^0.25
// (vp_set_density, :1272-1287), both formed in double and rounded as the
// host did; M = rho vp^2, MU = rho vs^2, BU = 1/rho in float.
__global__ void k_model_moduli(const float* __restrict__ vp, const float* __restrict__ vs,
                               const float* __restrict__ rho, int nx, int ny, int nz, int bl, int bf, int bu,
                               int nnx, int nny, int nnz, int px, float* __restrict__ M, float* __restrict__ MU,
                               float* __restrict__ BU) {
    const long total = (long)nnx * nny * nnz;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nnx);
        const long r = t / nnx;
        const int j = (int)(r % nny), k = (int)(r / nny);
        const int ix = min(max(i - bl, 0), nx - 1), jy = min(max(j - bf, 0), ny - 1), kz = min(max(k - bu, 0), nz - 1);
        const long q = ((long)ix * ny + jy) * nz + kz;
        const float a = vp[q];
        const float b = vs ? vs[q] : (float)((double)a / sqrt(3.0));
        const float rr = rho ? rho[q] : (float)(1000 * 0.23 * sqrt(sqrt((double)a)));
        const long n = ((long)k * nny + j) * px + i;
        M[n] = rr * a * a;
        MU[n] = rr * b * b;
        BU[n] = 1.0f / rr;
    }
}

// Wavefield snapshots (join_wavefield != 0: output_3d_wavefiled_{tao,vx,vz,
// vzp,vzs}, 3D_output_file.cu:2-151, gated at 3D_forward_or_back_together.cu:
// 50,198,341).  After the update of time index it with (it + 1) a multiple of
// wavefield_interval: ./wavefield/{name}-{it+1} (padded grid), the per-device
// file {name}-0-{it+1} (one device: the same grid) and cut-{name}-{it+1}
// (model region), raw float32 in disk order (z fastest), for name = tp, vx,
// vz, vzp, vzs (vzs = vz - vzp).  When the pass reaches its last step,
// ./wavefield is renamed ./wavefield1 (model pass), ./wavefield2 (source) or
// ./wavefield3 (receiver), replacing an older one.
int write_snapshots(erm_ctx* c, int label) {
    ::mkdir("wavefield", 0755);
    const erm_params& p = c->p;
    const char* names[5] = {"tp", "vx", "vz", "vzp", "vzs"};
    const int fid[5] = {ERM_F_TP, ERM_F_VX, ERM_F_VZ, ERM_F_VZP, ERM_F_VZS};
    const size_t N = (size_t)c->nnx * c->nny * c->nnz;
    std::vector<float> full(N), disk(N), cut((size_t)p.nx * p.ny * p.nz);
    for (int a = 0; a < 5; ++a) {
        int rc = erm_get_field(c, fid[a], full.data());
        if (rc) return rc;
        for (int ix = 0; ix < c->nnx; ++ix)
            for (int iy = 0; iy < c->nny; ++iy)
                for (int iz = 0; iz < c->nnz; ++iz)
                    disk[((size_t)ix * c->nny + iy) * c->nnz + iz] = full[((size_t)iz * c->nny + iy) * c->nnx + ix];
        for (int ix = 0; ix < p.nx; ++ix)
            for (int iy = 0; iy < p.ny; ++iy)
                for (int iz = 0; iz < p.nz; ++iz)
                    cut[((size_t)ix * p.ny + iy) * p.nz + iz] =
                        full[((size_t)(iz + p.bu) * c->nny + iy + p.bf) * c->nnx + ix + p.bl];
        const std::string f1 = format("wavefield/%s-%d", names[a], label), f2 = format("wavefield/%s-0-%d", names[a], label),
                          f3 = format("wavefield/cut-%s-%d", names[a], label);
        if (write_raw(f1.c_str(), disk.data(), N) || write_raw(f2.c_str(), disk.data(), N) ||
            write_raw(f3.c_str(), cut.data(), cut.size()))
            return c->fail(ERM_E_IO, "can not write wavefield snapshot %s", f1.c_str());
    }
    return 0;
}
int finish_snapshots(erm_ctx* c, int pass) {
    const std::string dst = format("wavefield%d", pass + 1);
    std::error_code ec;
    std::filesystem::remove_all(dst, ec);
    std::filesystem::rename("wavefield", dst, ec);
    if (ec) return c->fail(ERM_E_IO, "can not rename wavefield to %s: %s", dst.c_str(), ec.message().c_str());
    return 0;
}

int run_pass(erm_ctx* c, int pass, int it_begin, int nsteps, bool strips = false, bool hist = false) {
    if (pass < ERM_PASS_MODEL || pass > ERM_PASS_RECEIVER) return c->fail(ERM_E_PARAM, "unknown pass %d", pass);
    const int which = pass == ERM_PASS_MODEL ? 0 : 1;
    if (!c->have_model[which]) return c->fail(ERM_E_STATE, "%s model not loaded", which ? "smoothed" : "true");
    if (nsteps < 0) return c->fail(ERM_E_PARAM, "nsteps < 0");
    if (pass == ERM_PASS_RECEIVER) {
        if (!c->have_record) return c->fail(ERM_E_STATE, "no receiver record: run erm_forward_real or erm_set_record first");
        if (it_begin >= c->p.lt || it_begin - nsteps + 1 < 0) return c->fail(ERM_E_PARAM, "receiver pass time range out of [0, lt)");
    } else {
        if (it_begin < 0) return c->fail(ERM_E_PARAM, "it_begin < 0");
        if (pass == ERM_PASS_MODEL && it_begin + nsteps > c->p.lt) return c->fail(ERM_E_PARAM, "model pass beyond lt");
    }
    const long comp = (long)c->p.lt * c->nr;
    const int rb = 256, rg = (c->nr + rb - 1) / rb;
    if (c->profiling) {
        size_t need = (size_t)3 * nsteps;
        while (c->kev.size() < need) {
            hipEvent_t e;
            HIPC(c, hipEventCreate(&e));
            c->kev.push_back(e);
        }
    }
    const bool snap = c->p.join_wavefield != 0 && c->p.wavefield_interval > 0 && !strips;
    bool wrote = false;
    HIPC(c, hipEventRecord(c->pass_ev[0], c->stream));
    for (int s = 0; s < nsteps; ++s) {
        const int it = pass == ERM_PASS_RECEIVER ? it_begin - s : it_begin + s;
        if (pass == ERM_PASS_RECEIVER && c->nr > 0)
            hipLaunchKernelGGL(k_inject, dim3(rg), dim3(rb), 0, c->stream, c->f[0], c->f[1], c->f[2], c->rec_idx, c->nr,
                               c->rec, comp, (long)it * c->nr);
        if (c->profiling) HIPC(c, hipEventRecord(c->kev[3 * s], c->stream));
        do_stress(c, which, pass == ERM_PASS_RECEIVER ? -1L : c->src, it);
        if (c->profiling) HIPC(c, hipEventRecord(c->kev[3 * s + 1], c->stream));
        if (strips && pass == ERM_PASS_SOURCE) strip_stress(c, c->f, it, 0);
        do_velocity(c, which, pass == ERM_PASS_SOURCE ? MODE_EXCITE : pass == ERM_PASS_RECEIVER ? MODE_IMAGE : MODE_PLAIN, it);
        if (c->profiling) HIPC(c, hipEventRecord(c->kev[3 * s + 2], c->stream));
        if (pass == ERM_PASS_RECEIVER && c->poyn[0])
            hipLaunchKernelGGL(k_poynting, dim3((unsigned)std::min<size_t>((c->ncell + 255) / 256, 16384)), dim3(256), 0,
                               c->stream, make_wave(c), (const int*)c->ex_time, c->poyn[0], c->poyn[1], c->poyn[2],
                               c->poyn[3], c->poyn[4], c->poyn[5], (long)c->ncell, it);
        if (strips && pass == ERM_PASS_SOURCE) strip_velocity(c, c->f, it + 1, 0);
        if (hist && (it + 1) % c->hist_decim == 0) {  // a stored step of the decimated history
            const erm_params& P = c->p;
            const long nm = (long)P.nx * P.ny * P.nz;
            const float* slot = c->hist + (size_t)(it / c->hist_decim) * 4 * nm;
            const dim3 hg((unsigned)std::min<long>((nm + 255) / 256, 16384));
            if (pass == ERM_PASS_SOURCE)
                hipLaunchKernelGGL(k_hist_save, hg, dim3(256), 0, c->stream, make_wave(c), (float*)slot, P.nx, P.ny, P.nz,
                                   P.bl, P.bf, P.bu, c->px, c->pxy);
            else if (pass == ERM_PASS_RECEIVER)
                hipLaunchKernelGGL(k_xcorr_hist, hg, dim3(256), 0, c->stream, slot, make_wave(c), c->ximg[0], c->ximg[1],
                                   c->ximg[2], P.nx, P.ny, P.nz, P.bl, P.bf, P.bu, c->px, c->pxy, (float)c->hist_decim);
        }
        if (strips && pass == ERM_PASS_RECEIVER) {
            // source state {sigma^{it+1/2}, v^{it+1}}: complete the slab cells, image, step back
            strip_stress(c, c->fr, it, 1);
            const erm_params& P = c->p;
            const long nm = (long)P.nx * P.ny * P.nz;
            hipLaunchKernelGGL(k_xcorr, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, c->stream, wave_of(c->fr),
                               make_wave(c), c->ximg[0], c->ximg[1], c->ximg[2], P.nx, P.ny, P.nz, P.bl, P.bf, P.bu,
                               c->px, c->pxy);
            reverse_rest(c, it);
        }
        if (pass == ERM_PASS_MODEL && c->nr > 0)
            hipLaunchKernelGGL(k_record, dim3(rg), dim3(rb), 0, c->stream, c->f[0], c->f[1], c->f[2], c->rec_idx, c->nr,
                               c->rec, comp, (long)it * c->nr);
        if (snap && (it + 1) % c->p.wavefield_interval == 0) {
            HIPC(c, hipGetLastError());
            int rc = write_snapshots(c, it + 1);
            if (rc) return rc;
            wrote = true;
        }
        const bool last = pass == ERM_PASS_RECEIVER ? it == 0
                                                    : it + 1 == (pass == ERM_PASS_MODEL ? c->p.lt : c->p.lt_forward);
        if (wrote && last) {
            int rc = finish_snapshots(c, pass);
            if (rc) return rc;
            wrote = false;
        }
    }
    HIPC(c, hipGetLastError());
    HIPC(c, hipEventRecord(c->pass_ev[1], c->stream));
    HIPC(c, hipEventSynchronize(c->pass_ev[1]));
    float ms = 0;
    HIPC(c, hipEventElapsedTime(&ms, c->pass_ev[0], c->pass_ev[1]));
    c->last_pass_ms = ms;
    c->last_stress_ms = c->last_velocity_ms = 0;
    c->last_stress_n = c->last_velocity_n = 0;
    if (c->profiling) {
        for (int s = 0; s < nsteps; ++s) {
            float a = 0, b = 0;
            HIPC(c, hipEventElapsedTime(&a, c->kev[3 * s], c->kev[3 * s + 1]));
            HIPC(c, hipEventElapsedTime(&b, c->kev[3 * s + 1], c->kev[3 * s + 2]));
            c->last_stress_ms += a;
            c->last_velocity_ms += b;
        }
        c->last_stress_n = c->last_velocity_n = nsteps;
    }
    if (pass == ERM_PASS_MODEL) c->have_record = true;
    return 0;
}

}  // namespace

// Direct-wave mute of the modelled record (the reference calls
// cut_direct_shot_3D_surface_three_new after recording each step when vsp == 0,
// 3D_forward_or_back_together.cu:133-136; its body is not in the reference, so the
// window is defined here): at receiver r the samples
//   it < ceil(|x_r - x_s| / (vp_s dt) - 1e-6) + wl + cut_direct_wave
// of all three components are zeroed -- the direct P arrival from the source at
// the true model's P velocity at the source cell, the wavelet's duration (wl
// samples, the Ricker starts at it = 0) and cut_direct_wave samples of margin.
int mute_cut_samples(const erm_ctx* c, double vp_src, std::vector<int>& cut) {
    const erm_params& p = c->p;
    cut.assign(c->nr, 0);
    const double dts = p.dt / 1000.0;
    for (int iry = 0; iry < p.receiver_num_y; ++iry)
        for (int irx = 0; irx < p.receiver_num_x; ++irx) {
            const double ddx = (p.receiver_start_x + irx * p.receiver_interval_x - c->sxr) * (double)p.dx;
            const double ddy = (p.receiver_start_y + iry * p.receiver_interval_y - c->syr) * (double)p.dy;
            const double ddz = (p.receiver_start_z - c->szr) * (double)p.dz;
            const double t = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz) / (vp_src * dts);
            // 1e-6 sample of slack: a round t does not hinge on the last bit of vp_src
            const double n = std::ceil(t - 1e-6) + c->wl + p.cut_direct_wave;
            cut[(size_t)iry * p.receiver_num_x + irx] = (int)std::min<double>(n, p.lt);
        }
    return 0;
}

int mute_direct(erm_ctx* c) {
    if (c->p.cut_direct_wave <= 0 || c->p.vsp != 0 || c->nr <= 0) return 0;
    if (!c->have_model[0]) return c->fail(ERM_E_STATE, "true model not loaded");
    float mb[2];  // M = rho vp^2 and b = 1/rho at the source cell
    HIPC(c, hipMemcpyAsync(&mb[0], c->M[0] + c->src, 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipMemcpyAsync(&mb[1], c->BU[0] + c->src, 4, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    std::vector<int> cut;
    mute_cut_samples(c, std::sqrt((double)mb[0] * (double)mb[1]), cut);
    int* dcut = nullptr;
    if (dmalloc(c, &dcut, cut.size())) return ERM_E_NOMEM;
    hipError_t e = hipMemcpyAsync(dcut, cut.data(), cut.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
    const long n = 3L * c->p.lt * c->nr;
    if (e == hipSuccess)
        hipLaunchKernelGGL(k_mute, dim3((unsigned)std::min<long>((n + 255) / 256, 4096)), dim3(256), 0, c->stream,
                           c->rec, c->nr, (long)c->p.lt, (const int*)dcut);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dcut);
    if (e != hipSuccess) return c->fail(ERM_E_DEVICE, "direct-wave mute: %s", hipGetErrorString(e));
    return 0;
}

extern "C" {

const char* erm_last_error(const erm_ctx* ctx) { return ctx ? ctx->err.c_str() : global_error(); }
int erm_version(void) { return 100; }
size_t erm_params_size(void) { return sizeof(erm_params); }

int erm_device_count(int* n) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        set_global_error("hipGetDeviceCount: %s", hipGetErrorString(e));
        *n = 0;
        return ERM_E_DEVICE;
    }
    *n = c;
    return 0;
}

int erm_params_default(erm_params* p) {
    if (!p) return ERM_E_PARAM;
    params_default(p);
    return 0;
}

int erm_parse_par(int argc, const char* const* argv, erm_params* p) {
    if (!p) return ERM_E_PARAM;
    std::string err;
    int rc = parse_par(argc, argv, p, &err);
    if (rc) set_global_error("%s", err.c_str());
    return rc;
}

int erm_create(const erm_params* pin, int device, erm_ctx** out) {
    if (!pin || !out) { set_global_error("null argument"); return ERM_E_PARAM; }
    *out = nullptr;
    int R = pin->fd_order == 8 ? 4 : pin->fd_order == 12 ? 6 : -1;
    if (R < 0) { set_global_error("fd_order must be 8 or 12"); return ERM_E_PARAM; }
    std::string err;
    int rc = check_grid(pin, R, &err);
    if (rc) { set_global_error("%s", err.c_str()); return rc; }
    if (pin->poynting != 0 && pin->poynting != 1) { set_global_error("poynting must be 0 or 1"); return ERM_E_PARAM; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_global_error("no HIP device"); return ERM_E_DEVICE; }
    if (device < 0 || device >= ndev) { set_global_error("device %d out of range (%d devices)", device, ndev); return ERM_E_PARAM; }
    erm_ctx* c = new erm_ctx();
    c->p = *pin;
    if (c->p.lt_forward < 0) c->p.lt_forward = 2 * c->p.lt / 3;
    c->device = device;
    c->R = R;
    c->nnx = pin->nx + pin->bl + pin->br;
    c->nny = pin->ny + pin->bf + pin->bb;
    c->nnz = pin->nz + pin->bu + pin->bd;
    {
        const int align = pin->pitch_align == 1 ? 1 : (pin->pitch_align > 1 ? pin->pitch_align : 32);
        c->px = (c->nnx + align - 1) / align * align;  // 128-byte rows: a 64-wide tile row = 2 cache lines
    }
    c->xcd_map = pin->xcd_map == 1 || pin->xcd_map == 2 ? pin->xcd_map : 0;
    c->region_split = pin->cpml_split != 1;
    c->pxy = (long)c->px * c->nny;
    c->ncell = (size_t)c->pxy * c->nnz;
    for (int m = 0; m <= 6; ++m) c->coef.c[m] = 0.f;
    for (int m = 0; m <= R; ++m) c->coef.c[m] = (float)(R == 6 ? kCoeR6[m] : kCoeR4[m]);
    const double dts = pin->dt / 1000.0;
    c->cx = (float)(dts / pin->dx);
    c->cy = (float)(dts / pin->dy);
    c->cz = (float)(dts / pin->dz);
    // tile shapes 64 x {4, 8}; at radius 4 also 32 x {8, 16} (tile_x = 32)
    // default 32-wide at radius 4 (32 x 8 measured fastest with C-PML), 64-wide at radius 6
    c->tx = R == 4 && pin->tile_x != 64 ? 32 : 64;
    c->ty = c->tx == 64 ? (pin->tile_y == 4 ? 4 : 8) : (pin->tile_y == 16 ? 16 : 8);
    if (c->px % 32 == 0) {  // tiles over the updated columns only, their first column line-aligned
        c->x0 = R;
        c->xoff = (32 - R % 32) % 32;
    }
    {
        int tiles = ((c->nnx + c->tx - 1) / c->tx) * ((c->nny + c->ty - 1) / c->ty);
        int nzi = c->nnz - 2 * R;
        int chunks = std::max(1, (4096 + tiles - 1) / tiles);
        int kc = (nzi + chunks - 1) / chunks;
        kc = std::max(kc, std::min(nzi, 8));
        c->kchunk = pin->zchunk > 0 ? std::min(pin->zchunk, nzi) : kc;
        const int kmax = max_sweep(c);
        if (kmax == 0) {
            set_global_error("x-y plane of %ld cells too large: %d planes exceed the 4 GiB span of 32-bit offsets",
                             c->pxy, 2 * R + 1);
            delete c;
            return ERM_E_SIZE;
        }
        if (c->kchunk > kmax) {
            if (pin->zchunk > 0) {
                set_global_error("zchunk %d too large for %ld-cell planes (at most %d: 32-bit offsets)", pin->zchunk,
                                 c->pxy, kmax);
                delete c;
                return ERM_E_PARAM;
            }
            c->kchunk = kmax;
        }
    }
#define CK(x)                         \
    do {                              \
        int r_ = (x);                 \
        if (r_) {                     \
            set_global_error("%s", c->err.c_str()); \
            free_all(c);              \
            delete c;                 \
            return r_;                \
        }                             \
    } while (0)
    if (hipSetDevice(device) != hipSuccess) { set_global_error("hipSetDevice failed"); delete c; return ERM_E_DEVICE; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { set_global_error("stream create failed"); delete c; return ERM_E_DEVICE; }
    if (pin->region_streams == 1 &&
        (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
         hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess)) {
        set_global_error("second stream create failed"); free_all(c); delete c; return ERM_E_DEVICE;
    }
    for (auto& e : c->pass_ev)
        if (hipEventCreate(&e) != hipSuccess) { set_global_error("event create failed"); free_all(c); delete c; return ERM_E_DEVICE; }
    for (auto& p : c->f) CK(vmalloc(c, &p, c->ncell));
    CK(vmalloc(c, &c->exbuf, (size_t)EX_COUNT * c->ncell));
    c->ex_time = reinterpret_cast<int*>(c->exbuf + (size_t)EX_TIME * c->ncell);
    for (int q = 0; q < 5; ++q) c->ex[q] = c->exbuf + (size_t)(EX_AMP + q) * c->ncell;
    for (int q = 0; q < 3; ++q) c->img[q] = c->exbuf + (size_t)(EX_IMG_TP + q) * c->ncell;
    if (pin->boundary == ERM_BOUNDARY_CPML) {
        c->cpml = true;
        const int n3[3] = {c->nnx, c->nny, c->nnz};
        const int lo[3] = {pin->bl, pin->bf, pin->bu}, hi[3] = {pin->br, pin->bb, pin->bd};
        for (int a = 0; a < 3; ++a) {
            c->slab_lo[a] = lo[a];
            c->slab_hi[a] = hi[a];
            c->slab_w[a] = lo[a] + hi[a] + 2;
            if (c->slab_w[a] > n3[a]) {  // layers cover the axis: keep the slab map inside [0, n)
                c->slab_hi[a] = std::max(0, n3[a] - lo[a] - 2);
                c->slab_w[a] = lo[a] + c->slab_hi[a] + 2;
            }
            c->pml.lo[a] = c->slab_lo[a];
            c->pml.hi[a] = c->slab_hi[a];
            c->pml.w[a] = c->slab_w[a];
        }
        const size_t sz[3] = {(size_t)c->nnz * c->nny * c->slab_w[0], (size_t)c->nnz * c->slab_w[1] * c->nnx,
                              (size_t)c->slab_w[2] * c->nny * c->nnx};
        for (int a = 0; a < 3; ++a) c->pml.sz[a] = (unsigned)((sz[a] + 3) / 4 * 4);  // 16-B aligned blocks
        const size_t nstate = pml_floats(c->pml, c->nnx, c->nny, c->nnz);
        if (nstate * 4 >= (1UL << 32)) {
            set_global_error("C-PML memory variables exceed 4 GB (32-bit buffer offsets)");
            free_all(c); delete c; return ERM_E_SIZE;
        }
        CK(dmalloc(c, &c->psi_buf, nstate));
        c->pml.psi = c->psi_buf;
        c->vmax_loaded = 0.0;
        CK(build_cpml(c));
    }
    // damping, h = d*dt/2 per axis (3D_zzzzz:261-369)
    {
        std::vector<double> du, dd, dl, dr, df, db;
        double cA = pin->boundary == ERM_BOUNDARY_CPML ? 0.0 : pin->coe_attenuation;  // CPML replaces damping
        face_profile(pin->bu, cA, false, du); face_profile(pin->bd, cA, true, dd);
        face_profile(pin->bl, cA, false, dl); face_profile(pin->br, cA, true, dr);
        face_profile(pin->bf, cA, false, df); face_profile(pin->bb, cA, true, db);
        std::vector<float> ax(c->nnx, 0.f), ay(c->nny, 0.f), az(c->nnz, 0.f);
        for (int i = 0; i < pin->bl; ++i) ax[i] += (float)dl[i];
        for (int i = 0; i < pin->br; ++i) ax[pin->bl + pin->nx + i] += (float)dr[i];
        for (int j = 0; j < pin->bf; ++j) ay[j] += (float)df[j];
        for (int j = 0; j < pin->bb; ++j) ay[pin->bf + pin->ny + j] += (float)db[j];
        for (int k = 0; k < pin->bu; ++k) az[k] += (float)du[k];
        for (int k = 0; k < pin->bd; ++k) az[pin->bu + pin->nz + k] += (float)dd[k];
        // the kernel forms h = hx+hy+hz; h = (ax+ay+az)*dt/2 up to rounding
        const float half = (float)(dts / 2.0);
        for (auto* v : {&ax, &ay, &az}) for (auto& x : *v) x *= half;
        std::vector<float>* axes[3] = {&ax, &ay, &az};
        for (int a = 0; a < 3; ++a) {
            CK(dmalloc(c, &c->hxyz[a], axes[a]->size()));
            if (hipMemcpy(c->hxyz[a], axes[a]->data(), axes[a]->size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
                set_global_error("hipMemcpy damping failed"); free_all(c); delete c; return ERM_E_DEVICE;
            }
        }
    }
    c->wavelet_h = ricker(pin->freq, (float)dts);
    c->wl = (int)c->wavelet_h.size();
    CK(dmalloc(c, &c->wavelet, c->wavelet_h.size()));
    if (hipMemcpy(c->wavelet, c->wavelet_h.data(), c->wavelet_h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        set_global_error("hipMemcpy wavelet failed"); free_all(c); delete c; return ERM_E_DEVICE;
    }
    // receivers on the surface grid (write_or_add_shot_3D_surface_three, :127)
    c->nr = pin->receiver_num_x * pin->receiver_num_y;
    if (c->nr > 0) {
        std::vector<long> idx(c->nr);
        for (int iry = 0; iry < pin->receiver_num_y; ++iry)
            for (int irx = 0; irx < pin->receiver_num_x; ++irx) {
                int i = pin->bl + pin->receiver_start_x + irx * pin->receiver_interval_x;
                int j = pin->bf + pin->receiver_start_y + iry * pin->receiver_interval_y;
                int k = pin->bu + pin->receiver_start_z;
                if (i < 0 || i >= c->nnx || j < 0 || j >= c->nny || k < 0 || k >= c->nnz) {
                    set_global_error("receiver (%d,%d) outside the grid", irx, iry);
                    free_all(c); delete c; return ERM_E_PARAM;
                }
                idx[(size_t)iry * pin->receiver_num_x + irx] = (long)k * c->pxy + (long)j * c->px + i;
            }
        CK(dmalloc(c, &c->rec_idx, idx.size()));
        if (hipMemcpy(c->rec_idx, idx.data(), idx.size() * sizeof(long), hipMemcpyHostToDevice) != hipSuccess) {
            set_global_error("hipMemcpy receivers failed"); free_all(c); delete c; return ERM_E_DEVICE;
        }
        CK(dmalloc(c, &c->rec, (size_t)3 * pin->lt * c->nr));
    }
    if (pin->poynting == 1)
        for (auto& q : c->poyn) CK(vmalloc(c, &q, c->ncell));
    CK(erm_zero_wavefield(c));
    CK(erm_zero_excitation(c));
    for (auto& p : c->img) CK(fill(c, p, c->ncell, 0.f));
    CK(erm_set_shot(c, 0, 0, 0));
    if (hipStreamSynchronize(c->stream) != hipSuccess) { set_global_error("init sync failed"); free_all(c); delete c; return ERM_E_DEVICE; }
#undef CK
    *out = c;
    return 0;
}

int erm_destroy(erm_ctx* c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return 0;
}

int erm_get_dims(const erm_ctx* c, int* nnx, int* nny, int* nnz, int* wl) {
    if (!c) return ERM_E_PARAM;
    if (nnx) *nnx = c->nnx;
    if (nny) *nny = c->nny;
    if (nnz) *nnz = c->nnz;
    if (wl) *wl = c->wl;
    return 0;
}

int erm_load_model(erm_ctx* c, int which, const float* vp, const float* vs, const float* rho) {
    if (!c) return ERM_E_PARAM;
    if (which < 0 || which > 1 || !vp) return c->fail(ERM_E_PARAM, "erm_load_model: bad arguments");
    HIPC(c, hipSetDevice(c->device));
    const erm_params& p = c->p;
    const size_t n = (size_t)p.nx * p.ny * p.nz;
    // the padded cells replicate model cells, so the model region decides validity and vmax
    double vm = 0;
    for (size_t q = 0; q < n; ++q) {
        const double a = vp[q];
        const double b = vs ? (double)vs[q] : (double)(float)(a / std::sqrt(3.0));
        const double r = rho ? (double)rho[q] : (double)(float)(1000 * 0.23 * std::sqrt(std::sqrt(a)));
        if (!(r > 0) || !(a > 0) || b < 0)
            return c->fail(ERM_E_PARAM, "model has non-positive vp/rho or negative vs at model cell %zu", q);
        vm = std::max(vm, a);
    }
    if (!c->M[which]) {
        if (vmalloc(c, &c->M[which], c->ncell) || vmalloc(c, &c->MU[which], c->ncell) || vmalloc(c, &c->BU[which], c->ncell))
            return ERM_E_NOMEM;
    }
    float* raw = nullptr;  // model-region arrays in disk order, for the padding kernel
    if (dmalloc(c, &raw, 3 * n)) return ERM_E_NOMEM;
    const float* srcs[3] = {vp, vs, rho};
    hipError_t e = hipSuccess;
    for (int a = 0; a < 3 && e == hipSuccess; ++a)
        if (srcs[a]) e = hipMemcpyAsync(raw + a * n, srcs[a], n * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const long total = (long)c->nnx * c->nny * c->nnz;
        hipLaunchKernelGGL(k_model_moduli, dim3((unsigned)std::min<long>((total + 255) / 256, 65536)), dim3(256), 0, c->stream,
                           raw, vs ? raw + n : nullptr, rho ? raw + 2 * n : nullptr, p.nx, p.ny, p.nz, p.bl, p.bf, p.bu,
                           c->nnx, c->nny, c->nnz, c->px, c->M[which], c->MU[which], c->BU[which]);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(raw);
    if (e != hipSuccess) return c->fail(ERM_E_DEVICE, "erm_load_model: %s", hipGetErrorString(e));
    c->have_model[which] = true;
    c->vmax_loaded = std::max(c->vmax_loaded, vm);
    if (c->cpml) return build_cpml(c);
    return 0;
}

int erm_load_model_files(erm_ctx* c) {
    if (!c) return ERM_E_PARAM;
    const erm_params& p = c->p;
    const size_t n = (size_t)p.nx * p.ny * p.nz;
    std::vector<float> vp(n), vs(n), rho(n);
    for (int which = 0; which < 2; ++which) {
        const char* fv = which ? p.s_velocity : p.velocity;
        const char* fs = which ? p.s_velocity1 : p.velocity1;
        const char* fr = which ? p.s_density : p.density;
        if (which == 1 && (p.smooth_time_vp != 0 || !fv[0]))
            return c->fail(ERM_E_PARAM, "smoothed models must be supplied as files (smooth_time_*=0); "
                                        "the SU smooth3d path is not part of this library");
        if (read_raw(fv, vp.data(), n)) return c->fail(ERM_E_IO, "Can not open this file : %s", fv);
        bool use_vs = p.join_vs != 0, use_rho = p.join_den != 0;
        if (use_vs && read_raw(fs, vs.data(), n)) return c->fail(ERM_E_IO, "Can not open this file : %s", fs);
        if (use_rho && read_raw(fr, rho.data(), n)) return c->fail(ERM_E_IO, "Can not open this file : %s", fr);
        int rc = erm_load_model(c, which, vp.data(), use_vs ? vs.data() : nullptr, use_rho ? rho.data() : nullptr);
        if (rc) return rc;
    }
    return 0;
}

int erm_set_shot(erm_ctx* c, int isx, int isy, int isz) {
    if (!c) return ERM_E_PARAM;
    const erm_params& p = c->p;
    int sxr = p.shot_start_x + isx * p.shot_interval_x;
    int syr = p.shot_start_y + isy * p.shot_interval_y;
    int szr = p.shot_start_z + isz * p.shot_interval_z;
    int i = p.bl + sxr, j = p.bf + syr, k = p.bu + szr;
    if (i < c->R || i >= c->nnx - c->R || j < c->R || j >= c->nny - c->R || k < c->R || k >= c->nnz - c->R)
        return c->fail(ERM_E_PARAM, "source (%d,%d,%d) is not an updated cell of the padded grid", i, j, k);
    c->sxr = sxr; c->syr = syr; c->szr = szr;
    c->src = (long)k * c->pxy + (long)j * c->px + i;
    return 0;
}

int erm_zero_wavefield(erm_ctx* c) {
    if (!c) return ERM_E_PARAM;
    for (auto& p : c->f) if (fill(c, p, c->ncell, 0.f)) return ERM_E_DEVICE;
    if (c->cpml && fill(c, c->psi_buf, psi_total(c->pml), 0.f)) return ERM_E_DEVICE;
    return 0;
}

int erm_zero_excitation(erm_ctx* c) {
    if (!c) return ERM_E_PARAM;
    if (filli(c, c->ex_time, c->ncell, -1)) return ERM_E_DEVICE;
    for (auto& p : c->ex) if (fill(c, p, c->ncell, 0.f)) return ERM_E_DEVICE;
    for (auto& p : c->poyn) if (p && fill(c, p, c->ncell, 0.f)) return ERM_E_DEVICE;
    return 0;
}

int erm_rtm_step(erm_ctx* c, int pass, int it_begin, int nsteps) {
    if (!c) return ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    return run_pass(c, pass, it_begin, nsteps);
}

int erm_forward_real(erm_ctx* c, float* rec) {
    if (!c) return ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    int rc = erm_zero_wavefield(c);
    if (rc) return rc;
    if ((rc = run_pass(c, ERM_PASS_MODEL, 0, c->p.lt)) || (rc = mute_direct(c))) return rc;
    if (rec) return erm_get_record(c, rec);
    return 0;
}

int erm_forward_smoothed(erm_ctx* c) {
    if (!c) return ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    int rc;
    if ((rc = erm_zero_wavefield(c)) || (rc = erm_zero_excitation(c))) return rc;
    return run_pass(c, ERM_PASS_SOURCE, 0, c->p.lt_forward);
}

int erm_backward_smoothed(erm_ctx* c, const float* rec) {
    if (!c) return ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    int rc;
    if (rec && (rc = erm_set_record(c, rec))) return rc;
    if ((rc = erm_zero_wavefield(c))) return rc;
    for (auto& p : c->img) if (fill(c, p, c->ncell, 0.f)) return ERM_E_DEVICE;
    for (auto& p : c->poyn) if (p && fill(c, p, c->ncell, 0.f)) return ERM_E_DEVICE;
    return run_pass(c, ERM_PASS_RECEIVER, c->p.lt - 1, c->p.lt);
}

int erm_set_record(erm_ctx* c, const float* rec) {
    if (!c || !rec) return ERM_E_PARAM;
    if (c->nr == 0) { c->have_record = true; return 0; }
    HIPC(c, hipMemcpyAsync(c->rec, rec, (size_t)3 * c->p.lt * c->nr * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    c->have_record = true;
    return 0;
}

int erm_get_record(erm_ctx* c, float* rec) {
    if (!c || !rec) return ERM_E_PARAM;
    if (c->nr == 0) return 0;
    HIPC(c, hipMemcpyAsync(rec, c->rec, (size_t)3 * c->p.lt * c->nr * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}

int erm_get_field(erm_ctx* c, int field, float* dst) {
    if (!c || !dst) return ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    const size_t N = (size_t)c->nnx * c->nny * c->nnz;
    if (field >= 0 && field < 13) return download(c, c->f[field], dst);
    if (field >= ERM_F_VXS && field <= ERM_F_VZS) {
        int k = field - ERM_F_VXS;
        std::vector<float> vp(N);
        int rc = download(c, c->f[k], dst);
        if (!rc) rc = download(c, c->f[10 + k], vp.data());
        if (rc) return rc;
        for (size_t n = 0; n < N; ++n) dst[n] -= vp[n];
        return 0;
    }
    if (field == ERM_F_EX_TIME) {
        std::vector<int> t(c->ncell);
        HIPC(c, hipMemcpyAsync(t.data(), c->ex_time, c->ncell * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < c->nnz; ++k)
            for (int j = 0; j < c->nny; ++j)
                for (int i = 0; i < c->nnx; ++i)
                    dst[((size_t)k * c->nny + j) * c->nnx + i] = (float)t[(size_t)k * c->pxy + (size_t)j * c->px + i];
        return 0;
    }
    if (field >= ERM_F_EX_AMP && field <= ERM_F_EX_VZP) return download(c, c->ex[field - ERM_F_EX_AMP], dst);
    if (field >= ERM_F_IMG_TP && field <= ERM_F_IMG_PS) return download(c, c->img[field - ERM_F_IMG_TP], dst);
    if (field >= ERM_F_POYN_SX && field <= ERM_F_POYN_SZ) {  // -ex_tp ex_vP
        std::vector<float> tp(N);
        int rc = download(c, c->ex[1], tp.data());
        if (!rc) rc = download(c, c->ex[2 + field - ERM_F_POYN_SX], dst);
        if (rc) return rc;
        for (size_t n = 0; n < N; ++n) dst[n] = -tp[n] * dst[n];
        return 0;
    }
    if (field >= ERM_F_POYN_RPX && field <= ERM_F_POYN_RSZ) {
        if (!c->poyn[0]) return c->fail(ERM_E_STATE, "receiver Poynting vectors need poynting = 1 at erm_create");
        return download(c, c->poyn[field - ERM_F_POYN_RPX], dst);
    }
    if (field >= ERM_F_DIV && field <= ERM_F_CURLZ) {
        int rc = decompose(c, field - ERM_F_DIV);
        return rc ? rc : download(c, c->decbuf, dst);
    }
    return c->fail(ERM_E_PARAM, "unknown field %d", field);
}

int erm_set_field(erm_ctx* c, int field, const float* src) {
    if (!c || !src) return ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    if (field >= 0 && field < 13) return upload(c, c->f[field], src);
    return c->fail(ERM_E_PARAM, "field %d is not writable", field);
}

int erm_get_images(erm_ctx* c, float* tp, float* pp, float* ps) {
    if (!c) return ERM_E_PARAM;
    const erm_params& p = c->p;
    const size_t N = (size_t)c->nnx * c->nny * c->nnz;
    std::vector<float> full(N);
    float* outs[3] = {tp, pp, ps};
    for (int a = 0; a < 3; ++a) {
        if (!outs[a]) continue;
        int rc = erm_get_field(c, ERM_F_IMG_TP + a, full.data());
        if (rc) return rc;
        // output_file_xyz_boundary (3D_zzzzz:65-85): cut the model region, disk order
        for (int ix = 0; ix < p.nx; ++ix)
            for (int iy = 0; iy < p.ny; ++iy)
                for (int iz = 0; iz < p.nz; ++iz)
                    outs[a][((size_t)ix * p.ny + iy) * p.nz + iz] =
                        full[((size_t)(iz + p.bu) * c->nny + iy + p.bf) * c->nnx + ix + p.bl];
    }
    return 0;
}

static int write_cut(erm_ctx* c, int field, const std::string& path) {
    const erm_params& p = c->p;
    std::vector<float> full((size_t)c->nnx * c->nny * c->nnz), cut((size_t)p.nx * p.ny * p.nz);
    int rc = erm_get_field(c, field, full.data());
    if (rc) return rc;
    for (int ix = 0; ix < p.nx; ++ix)
        for (int iy = 0; iy < p.ny; ++iy)
            for (int iz = 0; iz < p.nz; ++iz)
                cut[((size_t)ix * p.ny + iy) * p.nz + iz] =
                    full[((size_t)(iz + p.bu) * c->nny + iy + p.bf) * c->nnx + ix + p.bl];
    if (write_raw(path.c_str(), cut.data(), cut.size())) return c->fail(ERM_E_IO, "can not open file : %s", path.c_str());
    return 0;
}

int erm_write_results(erm_ctx* c, const char* dir) {
    if (!c || !dir) return ERM_E_PARAM;
    const char* names[3] = {"tp", "pp", "ps"};
    for (int a = 0; a < 3; ++a) {
        std::string path = c->p.vsp == 0 ? format("%s/result-%s-%d-%d", dir, names[a], c->syr, c->sxr)
                                         : format("%s/result-%s-%d", dir, names[a], c->szr);
        int rc = write_cut(c, ERM_F_IMG_TP + a, path);
        if (rc) return rc;
    }
    return 0;
}

int erm_write_excitation(erm_ctx* c, const char* dir) {
    if (!c || !dir) return ERM_E_PARAM;
    const char* names[6] = {"time", "amp", "tp", "vxp", "vyp", "vzp"};
    for (int a = 0; a < 6; ++a) {
        int rc = write_cut(c, ERM_F_EX_TIME + a, format("%s/cut-ex-%s-%d-%d", dir, names[a], c->syr, c->sxr));
        if (rc) return rc;
    }
    return 0;
}

int erm_set_profiling(erm_ctx* c, int on) {
    if (!c) return ERM_E_PARAM;
    c->profiling = on != 0;
    return 0;
}

int erm_get_timing(erm_ctx* c, double* pass_ms, double* stress_ms, long* stress_n, double* vel_ms, long* vel_n) {
    if (!c) return ERM_E_PARAM;
    if (pass_ms) *pass_ms = c->last_pass_ms;
    if (stress_ms) *stress_ms = c->last_stress_ms;
    if (stress_n) *stress_n = c->last_stress_n;
    if (vel_ms) *vel_ms = c->last_velocity_ms;
    if (vel_n) *vel_n = c->last_velocity_n;
    return 0;
}

int erm_algorithmic_bytes(const erm_ctx* c, int pass, double* sb, double* vb) {
    if (!c) return ERM_E_PARAM;
    // stress: v(3) read + 7 stress/tp rw (14) + M, mu (2) = 19 floats
    // velocity: 7 stress/tp read + 6 v/vP rw (12) + 1/rho (1) = 20 floats
    //           + ex_amp (source pass) or ex_time (receiver pass)
    // + C-PML memory variables (8 B per psi at every updated cell inside its layer)
    if (sb) *sb = 19.0 * 4.0 + (c->cpml ? c->pml_extra_stress : 0.0);
    if (vb) *vb = 20.0 * 4.0 + ((pass == ERM_PASS_SOURCE || pass == ERM_PASS_RECEIVER) ? 4.0 : 0.0) +
                  (c->cpml ? c->pml_extra_velocity : 0.0);
    return 0;
}

static int build_strip_index(erm_ctx* c) {
    if (c->strip_idx) return 0;
    const erm_params& P = c->p;
    const int R = c->R;
    if (P.bl < R - 1 || P.br < R - 1 || P.bf < R - 1 || P.bb < R - 1 || P.bu < R - 1 || P.bd < R - 1)
        return c->fail(ERM_E_PARAM, "boundary strips need every boundary width >= %d", R - 1);
    if (c->ncell >= (1UL << 31)) return c->fail(ERM_E_SIZE, "grid too large for 32-bit strip indices");
    if (P.nx < 3 || P.ny < 3 || P.nz < 3) return c->fail(ERM_E_SIZE, "model region too small for reconstruction");
    // reconstructed region I: the model region shrunk by one cell on every face.
    // Strip cells: the R-cell face slabs around I (where the stencils of I
    // reach) and the rest of the model region's border shell (its edges and
    // corners), so every model-region cell is either reconstructed or restored.
    const int x0 = P.bl + 1, x1 = P.bl + P.nx - 1, y0 = P.bf + 1, y1 = P.bf + P.ny - 1, z0 = P.bu + 1, z1 = P.bu + P.nz - 1;
    std::vector<char> mark((size_t)c->nnx * c->nny * c->nnz, 0);
    auto set = [&](int i, int j, int k) { mark[((size_t)k * c->nny + j) * c->nnx + i] = 1; };
    for (int k = z0; k < z1; ++k)
        for (int j = y0; j < y1; ++j)
            for (int r = 1; r <= R; ++r) { set(x0 - r, j, k); set(x1 - 1 + r, j, k); }
    for (int k = z0; k < z1; ++k)
        for (int r = 1; r <= R; ++r)
            for (int i = x0; i < x1; ++i) { set(i, y0 - r, k); set(i, y1 - 1 + r, k); }
    for (int r = 1; r <= R; ++r)
        for (int j = y0; j < y1; ++j)
            for (int i = x0; i < x1; ++i) { set(i, j, z0 - r); set(i, j, z1 - 1 + r); }
    for (int k = P.bu; k < P.bu + P.nz; ++k)
        for (int j = P.bf; j < P.bf + P.ny; ++j)
            for (int i = P.bl; i < P.bl + P.nx; ++i)
                if (i < x0 || i >= x1 || j < y0 || j >= y1 || k < z0 || k >= z1) set(i, j, k);
    std::vector<int> idx;
    for (int k = 0; k < c->nnz; ++k)
        for (int j = 0; j < c->nny; ++j)
            for (int i = 0; i < c->nnx; ++i)
                if (mark[((size_t)k * c->nny + j) * c->nnx + i]) idx.push_back((int)((long)k * c->pxy + (long)j * c->px + i));
    c->n_strip = (int)idx.size();
    if (dmalloc(c, &c->strip_idx, idx.size())) return ERM_E_NOMEM;
    HIPC(c, hipMemcpy(c->strip_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

int erm_strip_bytes(const erm_ctx* c, double* bytes) {
    if (!c || !bytes) return ERM_E_PARAM;
    const erm_params& P = c->p;
    const double ax = P.nx - 2, ay = P.ny - 2, az = P.nz - 2;
    // face slabs around the shrunk region plus the model region's border edges (upper bound)
    const double n = 2.0 * c->R * (ay * az + ax * az + ax * ay) + 4.0 * ((double)P.nx + P.ny + P.nz);
    *bytes = 4.0 * kStripFields * n * (c->p.lt_forward + 1);
    return 0;
}

int erm_forward_smoothed_strips(erm_ctx* c) {
    if (!c) return ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    if (c->p.lt_forward != c->p.lt) return c->fail(ERM_E_PARAM, "reconstruction needs lt_forward == lt");
    int rc;
    if ((rc = build_strip_index(c))) return rc;
    const long need = (long)kStripFields * c->n_strip * (c->p.lt + 1);
    if (c->strip_steps != c->p.lt) {
        if (c->strips) (void)hipFree(c->strips);
        c->strips = nullptr;
        c->strip_steps = 0;
        size_t fr = 0, tot = 0;
        HIPC(c, hipMemGetInfo(&fr, &tot));
        if ((size_t)need * 4 + (64UL << 20) > fr)
            return c->fail(ERM_E_NOMEM, "boundary strips need %.1f GB, %.1f GB free", need * 4e-9, fr * 1e-9);
        if (dmalloc(c, &c->strips, (size_t)need)) return ERM_E_NOMEM;
        c->strip_steps = c->p.lt;
    }
    if (!c->wavelet_neg) {
        std::vector<float> wn(c->wavelet_h.size());
        for (size_t q = 0; q < wn.size(); ++q) wn[q] = -c->wavelet_h[q];
        if (dmalloc(c, &c->wavelet_neg, wn.size())) return ERM_E_NOMEM;
        HIPC(c, hipMemcpy(c->wavelet_neg, wn.data(), wn.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    for (auto& q : c->fr) if (!q && vmalloc(c, &q, c->ncell)) return ERM_E_NOMEM;
    if ((rc = fill(c, c->strips, (size_t)kStripFields * c->n_strip, 0.f))) return rc;  // slot 0: the zero initial state
    if ((rc = erm_zero_wavefield(c)) || (rc = erm_zero_excitation(c))) return rc;
    c->have_strips = false;
    if ((rc = run_pass(c, ERM_PASS_SOURCE, 0, c->p.lt, true))) return rc;
    for (int q = 0; q < 13; ++q)
        HIPC(c, hipMemcpyAsync(c->fr[q], c->f[q], c->ncell * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    c->have_strips = true;
    return 0;
}

int erm_reconstruct(erm_ctx* c, int it_end) {
    if (!c) return ERM_E_PARAM;
    if (!c->have_strips) return c->fail(ERM_E_STATE, "no boundary strips: run erm_forward_smoothed_strips first");
    if (it_end < 0 || it_end > c->p.lt) return c->fail(ERM_E_PARAM, "it_end out of [0, lt]");
    HIPC(c, hipSetDevice(c->device));
    for (int it = c->p.lt - 1; it >= it_end; --it) reverse_step(c, it);
    // complete the slab cells at level it_end - 1/2 (v, vP are already in place)
    if (it_end > 0) strip_stress(c, c->fr, it_end - 1, 1);
    else for (int q = 3; q <= 9; ++q) if (int rc = fill(c, c->fr[q], c->ncell, 0.f)) return rc;
    HIPC(c, hipGetLastError());
    HIPC(c, hipStreamSynchronize(c->stream));
    c->have_strips = false;  // the reconstruction wavefield has moved off the final state
    return 0;
}

int erm_get_reconstructed_field(erm_ctx* c, int field, float* dst) {
    if (!c || !dst || field < 0 || field >= 13 || !c->fr[field]) return c ? c->fail(ERM_E_PARAM, "bad field") : ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    return download(c, c->fr[field], dst);
}

int erm_history_bytes(const erm_ctx* c, int decim, double* bytes) {
    if (!c || !bytes || decim < 1) return ERM_E_PARAM;
    const erm_params& p = c->p;
    *bytes = 4.0 * 4.0 * (double)p.nx * p.ny * p.nz * (double)(p.lt / decim);
    return 0;
}

int erm_forward_smoothed_history(erm_ctx* c, int decim) {
    if (!c) return ERM_E_PARAM;
    if (decim < 1) return c->fail(ERM_E_PARAM, "decim must be >= 1");
    HIPC(c, hipSetDevice(c->device));
    if (c->p.lt_forward != c->p.lt) return c->fail(ERM_E_PARAM, "the source history needs lt_forward == lt");
    const size_t nm = (size_t)c->p.nx * c->p.ny * c->p.nz, need = (size_t)(c->p.lt / decim) * 4 * nm;
    if (need == 0) return c->fail(ERM_E_PARAM, "decim larger than lt");
    if (c->hist_n < need) {
        if (c->hist) (void)hipFree(c->hist);
        c->hist = nullptr;
        c->hist_n = 0;
        size_t fr = 0, tot = 0;
        HIPC(c, hipMemGetInfo(&fr, &tot));
        if (need * 4 + (64UL << 20) > fr)
            return c->fail(ERM_E_NOMEM, "the source history needs %.1f GB, %.1f GB free (raise decim)", need * 4e-9, fr * 1e-9);
        if (dmalloc(c, &c->hist, need)) return ERM_E_NOMEM;
        c->hist_n = need;
    }
    c->hist_decim = decim;
    c->have_hist = false;
    int rc;
    if ((rc = erm_zero_wavefield(c)) || (rc = erm_zero_excitation(c))) return rc;
    if ((rc = run_pass(c, ERM_PASS_SOURCE, 0, c->p.lt, false, true))) return rc;
    c->have_hist = true;
    return 0;
}

int erm_backward_smoothed_xcorr_history(erm_ctx* c, const float* rec) {
    if (!c) return ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    if (!c->have_hist) return c->fail(ERM_E_STATE, "no source history: run erm_forward_smoothed_history first");
    int rc;
    if (rec && (rc = erm_set_record(c, rec))) return rc;
    for (auto& q : c->ximg) if (!q && vmalloc(c, &q, c->ncell)) return ERM_E_NOMEM;
    if ((rc = erm_zero_wavefield(c))) return rc;
    for (auto& q : c->img) if ((rc = fill(c, q, c->ncell, 0.f))) return rc;
    for (auto& q : c->ximg) if ((rc = fill(c, q, c->ncell, 0.f))) return rc;
    return run_pass(c, ERM_PASS_RECEIVER, c->p.lt - 1, c->p.lt, false, true);
}

int erm_backward_smoothed_xcorr(erm_ctx* c, const float* rec) {
    if (!c) return ERM_E_PARAM;
    HIPC(c, hipSetDevice(c->device));
    if (!c->have_strips) return c->fail(ERM_E_STATE, "no boundary strips: run erm_forward_smoothed_strips first");
    int rc;
    if (rec && (rc = erm_set_record(c, rec))) return rc;
    for (auto& q : c->ximg) if (!q && vmalloc(c, &q, c->ncell)) return ERM_E_NOMEM;
    if ((rc = erm_zero_wavefield(c))) return rc;
    for (auto& q : c->img) if ((rc = fill(c, q, c->ncell, 0.f))) return rc;
    for (auto& q : c->ximg) if ((rc = fill(c, q, c->ncell, 0.f))) return rc;
    rc = run_pass(c, ERM_PASS_RECEIVER, c->p.lt - 1, c->p.lt, true);
    c->have_strips = false;
    return rc;
}

int erm_get_xcorr_images(erm_ctx* c, float* tp, float* pp, float* ps) {
    if (!c) return ERM_E_PARAM;
    if (!c->ximg[0]) return c->fail(ERM_E_STATE, "no cross-correlation images: run erm_backward_smoothed_xcorr first");
    const erm_params& p = c->p;
    const size_t N = (size_t)c->nnx * c->nny * c->nnz;
    std::vector<float> full(N);
    float* outs[3] = {tp, pp, ps};
    for (int a = 0; a < 3; ++a) {
        if (!outs[a]) continue;
        int rc = download(c, c->ximg[a], full.data());
        if (rc) return rc;
        for (int ix = 0; ix < p.nx; ++ix)
            for (int iy = 0; iy < p.ny; ++iy)
                for (int iz = 0; iz < p.nz; ++iz)
                    outs[a][((size_t)ix * p.ny + iy) * p.nz + iz] =
                        full[((size_t)(iz + p.bu) * c->nny + iy + p.bf) * c->nnx + ix + p.bl];
    }
    return 0;
}

int erm_synchronize(erm_ctx* c) {
    if (!c) return ERM_E_PARAM;
    HIPC(c, hipStreamSynchronize(c->stream));
    return 0;
}

int erm_autotune_shape(erm_ctx* c, int steps, int* tile_x, int* tile_y, int* zchunk, double* ms_per_step) {
    if (!c || steps <= 0) return c ? c->fail(ERM_E_PARAM, "steps must be positive") : ERM_E_PARAM;
    if (!c->have_model[1]) return c->fail(ERM_E_STATE, "smoothed model not loaded");
    HIPC(c, hipSetDevice(c->device));
    const bool prof = c->profiling;
    c->profiling = false;
    const int nzi = c->nnz - 2 * c->R;
    double best = 1e300;
    int bx = c->tx, bt = c->ty, bk = c->kchunk;
    // tile shapes: 64 x {4, 8}, and at radius 4 32 x {8, 16}
    std::vector<std::pair<int, int>> shapes = {{64, 4}, {64, 8}};
    if (c->R == 4) { shapes.push_back({32, 8}); shapes.push_back({32, 16}); }
    const int kcs[8] = {4, 5, 6, 7, 8, 12, 16, 32};  // z sweep: parallelism vs re-read z halos
    int rc = 0;
    for (const auto& sh : shapes)
        for (int kc : kcs) {
            if (kc > nzi && kc != kcs[0]) continue;
            if (!sweep_fits(c, std::min(kc, nzi))) continue;
            c->tx = sh.first;
            c->ty = sh.second;
            c->kchunk = std::min(kc, nzi);
            if ((rc = erm_zero_wavefield(c)) || (rc = erm_zero_excitation(c))) break;
            if ((rc = run_pass(c, ERM_PASS_SOURCE, 0, 1))) break;  // warm-up
            if ((rc = run_pass(c, ERM_PASS_SOURCE, 1, steps))) break;
            const double per = c->last_pass_ms / steps;
            if (per < best) { best = per; bx = c->tx; bt = c->ty; bk = c->kchunk; }
        }
    if (!rc && best == 1e300) {  // no candidate fits the 32-bit span: time the current shape
        c->tx = bx;
        c->ty = bt;
        c->kchunk = bk;
        if (!(rc = erm_zero_wavefield(c)) && !(rc = erm_zero_excitation(c)) && !(rc = run_pass(c, ERM_PASS_SOURCE, 0, 1)) &&
            !(rc = run_pass(c, ERM_PASS_SOURCE, 1, steps)))
            best = c->last_pass_ms / steps;
    }
    c->tx = bx;
    c->ty = bt;
    c->kchunk = bk;
    c->profiling = prof;
    if (rc) return rc;
    if ((rc = erm_zero_wavefield(c)) || (rc = erm_zero_excitation(c))) return rc;
    if (tile_x) *tile_x = bx;
    if (tile_y) *tile_y = bt;
    if (zchunk) *zchunk = bk;
    if (ms_per_step) *ms_per_step = best;
    return 0;
}

int erm_autotune(erm_ctx* c, int steps, int* tile_y, int* zchunk, double* ms_per_step) {
    return erm_autotune_shape(c, steps, nullptr, tile_y, zchunk, ms_per_step);
}

int erm_copy_bandwidth(int device, size_t bytes, int reps, double* ms_per_copy) {
    if (bytes < 16 || reps <= 0 || !ms_per_copy) { set_global_error("erm_copy_bandwidth: bad arguments"); return ERM_E_PARAM; }
    if (hipSetDevice(device) != hipSuccess) { set_global_error("hipSetDevice failed"); return ERM_E_DEVICE; }
    const long n4 = (long)(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, n4 * 16) != hipSuccess || hipMalloc((void**)&b, n4 * 16) != hipSuccess) {
        if (a) (void)hipFree(a);
        set_global_error("hipMalloc failed");
        return ERM_E_NOMEM;
    }
    (void)hipMemset(a, 0, n4 * 16);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // the grid-stride copy's rate depends on the grid (DRAM page locality of the
    // concurrently streamed window vs latency hiding): report the best of a sweep
    hipError_t e = hipSuccess;
    double best = 0;
    for (int blocks = 1024; blocks <= 16384 && e == hipSuccess; blocks *= 2) {
        hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n4);  // warm-up
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n4);
        (void)hipEventRecord(e1, 0);
        e = hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (best == 0 || ms / reps < best) best = ms / reps;
    }
    *ms_per_copy = best;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) { set_global_error("copy kernel failed: %s", hipGetErrorString(e)); return ERM_E_DEVICE; }
    return 0;
}

int erm_make_model3d(int kind, int nx, int ny, int nz, float dx, float dy, float dz, uint64_t seed, float* vp,
                     float* vs, float* rho) {
    int rc = make_model3d(kind, nx, ny, nz, dx, dy, dz, (unsigned long long)seed, vp, vs, rho);
    if (rc) set_global_error("erm_make_model3d: bad arguments");
    return rc;
}

int erm_write_raw(const char* path, const float* data, size_t n) {
    int rc = write_raw(path, data, n);
    if (rc) set_global_error("can not open file : %s", path ? path : "(null)");
    return rc;
}

int erm_read_raw(const char* path, float* data, size_t n) {
    int rc = read_raw(path, data, n);
    if (rc) set_global_error("can not open file : %s", path ? path : "(null)");
    return rc;
}

}  // extern "C"
