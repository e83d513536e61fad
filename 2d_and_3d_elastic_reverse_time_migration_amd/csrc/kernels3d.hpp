// kernels3d.hpp - gfx950 kernels of the 3D elastic RTM time step.
//
// Two fused kernels per time step replace the reference's five stencil
// launches (fwd_txxzzxzpp_3D, fwd_vx_3D, fwd_vy_3D, fwd_vz_3D,
// fwd_vxp_vzp_3D; 3D_forward_or_back_together.cu:83,100-104,118) plus the
// excitation (:282) and imaging (:442-446) passes:
//
//   k_stress   : sxx syy szz sxy sxz syz tp  <- v          (+ explosive source)
//   k_velocity : vx vy vz vxp vyp vzp        <- stress, tp (+ excitation | imaging)
//
// Layout: every volume is x-fastest, idx = k*pxy + j*px + i over the padded
// grid (3D_zzzzz:21).  A workgroup owns a TX x TY column tile (TX = 64 = one
// wavefront per row, coalesced along x) and sweeps a chunk of z planes: the
// current plane and its x/y halos are staged in LDS, the z-neighbourhood of
// each column lives in a register queue of 2R+1 planes, so every value of a
// stencil input is read from HBM about once per launch (halo re-reads hit
// L2/MALL).  Updates are in place: a stress cell reads only its own old
// stress and the (read-only) velocities, and vice versa, so the reference's
// "1"/"2" double buffers and their swap (exchange_wavefiled_new, :149) are
// not needed.
//
// Only cells with R <= i < nnx-R (same in y and z) are updated; the outer R
// cells of the padded grid stay zero, as in the 2D reference kernels that
// offset every index by `radius` (viscoelastic_equation.cu:90-92).
//
// Staggering (2D positions of viscoelastic_equation.cu:137-176 extended to y):
//   sxx syy szz tp model : (i, j, k)        vx : (i-1/2, j, k)
//   vy : (i, j-1/2, k)   vz : (i, j, k+1/2) sxy: (i-1/2, j-1/2, k)
//   sxz: (i-1/2, j, k+1/2)                  syz: (i, j-1/2, k+1/2)
//   D+ f[i] = sum_m c_m (f[i+m] - f[i-m+1]),  D- f[i] = sum_m c_m (f[i+m-1] - f[i-m]).
//
// Damping (Crank-Nicolson form of viscoelastic_equation.cu:151-180):
//   u' = [(1 - h) u + dt*rhs] / (1 + h),  h = d*dt/2,  d = dx(i)+dy(j)+dz(k)
// where the per-axis damping sums come from the 1-cos face profiles of
// coe_attenuation_3d_new2 (3D_zzzzz:261-369); the kernels take h per axis.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

namespace erm {

struct Coef {
    float c[7];  // c[1..R]
};

// Geometry shared by both kernels.
struct Geo {
    int nnx, nny, nnz;
    int px;          // row pitch (floats)
    long pxy;        // plane pitch (floats)
    float cx, cy, cz;  // dt/dx, dt/dy, dt/dz (dt in seconds)
    const float* hx;   // h = d*dt/2 per axis
    const float* hy;
    const float* hz;
    int kchunk;        // z planes per workgroup
    int z0;            // first chunk boundary, R - kchunk < z0 <= R (chunk_span)
    int x0;            // x of the first tile column (R with 32-wide tiles, which start at the first updated cell)
    int ntx, nty;      // column tiles in x and y
    int ntiles;        // ntx * nty * z chunks
    int per_xcd;       // tiles per XCD when xcd_map != 0
    int xcd_map;       // XCD-aware workgroup -> tile mapping
};

// Workgroup -> (x tile, y tile, z chunk).  The grid is 1-D, tile ids x
// fastest.  Workgroups are dispatched round-robin over the 8 XCDs (workgroup
// b runs on XCD b % 8), each with its own L2; with xcd_map the tiles are
// dealt so that XCD x walks a contiguous run [x*per, (x+1)*per) of tile ids
// and its tiles in flight are spatial neighbours (xcd_map 1; the default
// y-band order of the region-split lists is applied to the tile lists on the
// host instead).
// tile-list entries: tile id in bits 0..27, the C-PML axis mask in 28..30
constexpr int kTileIdMask = (1 << 28) - 1;

struct TileId {
    int bx, by, bz;
    bool valid;
};

// Updated planes [kb, ke) of z chunk bz: the chunk boundaries sit at
// z0 + m*kchunk, so with z0 < R the first chunk is shorter (empty when kb >= ke).
// The host places z0 so that boundaries fall on C-PML slab edges.
struct ZSpan {
    int kb, ke;
};
__host__ __device__ __forceinline__ ZSpan chunk_span(const Geo& g, int R, int bz) {
    const int a = g.z0 + bz * g.kchunk, b = a + g.kchunk, top = g.nnz - R;
    return ZSpan{a > R ? a : R, b < top ? b : top};
}
__device__ __forceinline__ TileId tile_of_block(const Geo& g, int blk) {
    int l = blk;
    if (g.xcd_map) l = (l & 7) * g.per_xcd + (l >> 3);
    TileId t;
    t.valid = l < g.ntiles;
    t.bx = l % g.ntx;
    l /= g.ntx;
    t.by = l % g.nty;
    t.bz = l / g.nty;
    return t;
}

struct Wave {  // the 13 propagated fields
    float *vx, *vy, *vz, *txx, *tyy, *tzz, *txy, *txz, *tyz, *tp, *vxp, *vyp, *vzp;
};

// Excitation state and images: nine arrays of n cells in one allocation, in
// the order of the EX_ ids (time holds ints).  One base pointer keeps the
// velocity kernels' rarely taken excitation/imaging branches from holding
// nine pointers in scalar registers across the z sweep.
struct Excite {
    float* base;
    long n;
};
enum { EX_TIME = 0, EX_AMP, EX_TP, EX_VXP, EX_VYP, EX_VZP, EX_IMG_TP, EX_IMG_PP, EX_IMG_PS, EX_COUNT };
__host__ __device__ __forceinline__ float* ex_array(const Excite& E, int q) { return E.base + q * E.n; }

enum { MODE_PLAIN = 0, MODE_EXCITE = 1, MODE_IMAGE = 2 };

// C-PML state (no reference counterpart; the reference damps instead).
// Coefficients a, b per axis at integer ("full") and staggered ("half":
// x-1/2, y-1/2, z+1/2) positions, interleaved per index in one array: axis a
// starts at float 4*(n_0 + .. + n_{a-1}) and holds n_a quadruples
// (a_full, b_full, a_half, b_half), so one 16-byte load fetches a position's
// four coefficients.
// 21 memory variables stored only in the absorbing slabs of their
// derivative's axis: along an axis of n cells the slab is [0, lo+1) U
// [n-hi-1, n) (width w = lo+hi+2, covering the half positions too).  Slab
// cell layouts: x slabs [nnz][nny][wx], y slabs [nnz][wy][nnx], z slabs
// [wz][nny][nnx]; sz[a] = slab cells rounded up to a multiple of 4.  The
// variables of one slab cell are interleaved: per axis a stress-side block
// of sz[a] triples at psi_stress(P, a) (x: exx, exy, exz; y: eyy, eyx, eyz;
// z: ezz, ezx, ezy) followed by a velocity-side block of sz[a] quadruples at
// psi_velocity(P, a) (x: sxx, sxy, sxz, tp; y: sxy, syy, syz, tp; z: sxz,
// syz, szz, tp).  The coefficients follow the memory variables in the same
// buffer (at float psi_total).  One load and one store per axis and cell,
// and one base pointer for the whole state, keep the C-PML kernels'
// scalar registers from spilling.
struct Pml {
    float* psi;
    int lo[3], hi[3], w[3];
    unsigned sz[3];
};
// float offsets of the stress-side and velocity-side blocks of axis a
__host__ __device__ __forceinline__ unsigned psi_stress(const Pml& P, int a) {
    return a == 0 ? 0u : a == 1 ? 7u * P.sz[0] : 7u * (P.sz[0] + P.sz[1]);
}
__host__ __device__ __forceinline__ unsigned psi_velocity(const Pml& P, int a) {
    return psi_stress(P, a) + 3u * P.sz[a];
}
__host__ __device__ __forceinline__ unsigned psi_total(const Pml& P) { return 7u * (P.sz[0] + P.sz[1] + P.sz[2]); }
// float offset of axis a's coefficient quadruples (in the coefficient array)
__host__ __device__ __forceinline__ unsigned coef_base(int nnx, int nny, int a) {
    return a == 0 ? 0u : a == 1 ? 4u * nnx : 4u * (nnx + nny);
}
// the same, in the state buffer
__host__ __device__ __forceinline__ unsigned pml_coef(const Pml& P, int nnx, int nny, int a) {
    return psi_total(P) + coef_base(nnx, nny, a);
}
__host__ __device__ __forceinline__ size_t pml_floats(const Pml& P, int nnx, int nny, int nnz) {
    return (size_t)psi_total(P) + 4 * (size_t)(nnx + nny + nnz);
}

__device__ __forceinline__ int slab_of(int i, int n, int lo, int hi) {
    return i <= lo ? i : (i >= n - hi - 1 ? lo + 1 + i - (n - hi - 1) : -1);
}
// e += (psi = b psi + a e).  Where a position's profile is zero (a = 0)
// psi stays at its initial 0 and e passes unchanged, so every slab cell can
// be updated and stored unconditionally.
__device__ __forceinline__ float pml_upd(float a, float b, float e, float& psi) {
    psi = b * psi + a * e;
    return e + psi;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Global accesses: a wave-uniform handle to an array at the first plane a
// workgroup touches, plus 32-bit byte offsets: per lane (in-plane, loop
// invariant) and per plane (uniform).  Two handle types with one interface:
//  - Rsrc, a raw buffer resource (four SGPRs; the per-plane offset goes in
//    the instruction's SGPR offset, so it costs no VGPRs).  Used by the tiles
//    without C-PML memory variables, which have scalar registers to spare and
//    keep their VGPR count, hence occupancy, lower this way.
//  - Base, a plain pointer (two SGPRs) addressed as base + zext(offset),
//    which compiles to global_load/store with a scalar base and a 32-bit
//    vector offset.  Used by the C-PML tiles: they touch 15 arrays, whose
//    buffer resources alone would take 60 of the 102 SGPRs and spill.
// Every access is in range by construction (halo indices are clamped, z
// planes stay within [0, nnz)).
struct Rsrc {
    __amdgpu_buffer_rsrc_t r;
};
struct Base {
    char* p;
};
template <class A>
__device__ __forceinline__ A mkacc(const void* base, long bytes);
template <>
__device__ __forceinline__ Rsrc mkacc<Rsrc>(const void* base, long bytes) {
    const unsigned n = bytes >= 0xFFFFFFFFL ? 0xFFFFFFFFu : (unsigned)bytes;
    // dword3 = 0x00020000: DATA_FORMAT = 32 bit, as required by offen raw loads on gfx9
    return Rsrc{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)n, 0x00020000)};
}
template <>
__device__ __forceinline__ Base mkacc<Base>(const void* base, long) {
    return Base{(char*)const_cast<void*>(base)};
}
// Streaming cache hint (gfx950 "nt", buffer cache-policy bit 1) for the
// accesses nothing else in the launch touches again: the updated-field
// stores, the cell's own-state loads (bldo) and the C-PML state.  Keeps L2
// for the stencil inputs that neighbouring tiles re-read (measured +1% on the
// 296x296x168 C-PML workload).
constexpr int kNtStore = 2, kNtOwn = 2;
__device__ __forceinline__ float bld(Rsrc r, unsigned vo, unsigned so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r.r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ float bldo(Rsrc r, unsigned vo, unsigned so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r.r, (int)vo, (int)so, kNtOwn));
}
__device__ __forceinline__ void bst(Rsrc r, unsigned vo, unsigned so, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r.r, (int)vo, (int)so, kNtStore);
}
__device__ __forceinline__ int bldi(Rsrc r, unsigned vo, unsigned so) {
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r.r, (int)vo, (int)so, 0);
}
__device__ __forceinline__ float bld(Base r, unsigned vo, unsigned so) {
    return *reinterpret_cast<const float*>(r.p + (vo + so));
}
__device__ __forceinline__ float bldo(Base r, unsigned vo, unsigned so) {
    const float* q = reinterpret_cast<const float*>(r.p + (vo + so));
    if constexpr (kNtOwn != 0) return __builtin_nontemporal_load(q);
    else return *q;
}
__device__ __forceinline__ void bst(Base r, unsigned vo, unsigned so, float v) {
    float* q = reinterpret_cast<float*>(r.p + (vo + so));
    if constexpr (kNtStore != 0) __builtin_nontemporal_store(v, q);
    else *q = v;
}
__device__ __forceinline__ int bldi(Base r, unsigned vo, unsigned so) {
    return *reinterpret_cast<const int*>(r.p + (vo + so));
}
// N consecutive floats per lane (C-PML state, Base handles only; the
// compiler merges them into one dwordx3 / dwordx4 access)
template <int N>
__device__ __forceinline__ void bldn(Base r, unsigned vo, unsigned so, float* v) {
    const float* q = reinterpret_cast<const float*>(r.p + (vo + so));
#pragma unroll
    for (int c = 0; c < N; ++c) {
        if constexpr (kNtOwn != 0) v[c] = __builtin_nontemporal_load(q + c);
        else v[c] = q[c];
    }
}
template <int N>
__device__ __forceinline__ void bstn(Base r, unsigned vo, unsigned so, const float* v) {
    float* q = reinterpret_cast<float*>(r.p + (vo + so));
#pragma unroll
    for (int c = 0; c < N; ++c) {
        if constexpr (kNtOwn != 0) __builtin_nontemporal_store(v[c], q + c);
        else q[c] = v[c];
    }
}

// ---------------------------------------------------------------------------
// k_stress: fwd_txxzzxzpp_3D (launch 3D_forward_or_back_together.cu:83) fused
// with add_source_3D (:36-38): the thread owning the source cell
// (plane src_k, in-plane offset src_col) adds wavelet[it] to sxx, syy, szz
// before updating them.
// HBM per cell-update: reads v(3), rw stress+tp(14), reads M, mu(2) = 76 B.
// ---------------------------------------------------------------------------
// Body of one workgroup's tile (block id `blk` of the shot's tiling), shared by
// the one-shot launch and the batched launch.
template <int R, int TX, int TY, int AX>
__device__ __forceinline__ void stress_tile(const Wave& w, const float* __restrict__ Mm,
                                            const float* __restrict__ MUm, const Geo& g, const Coef& cf,
                                            int src_k, int src_col, const float* __restrict__ wavelet,
                                            int wl, int it, const Pml& P, int blk, int tm) {
    constexpr int LX = TX + 2 * R, LY = TY + 2 * R;
    constexpr int NYH = (2 * R + TY - 1) / TY;
    __shared__ float sx[LY][LX];
    __shared__ float sy[LY][LX];
    __shared__ float sz[LY][LX];

    const TileId tid = tile_of_block(g, blk);
    if (!tid.valid) return;  // uniform per block
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = g.x0 + tid.bx * TX, j0 = tid.by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const ZSpan zs = chunk_span(g, R, tid.bz);
    const int kb = zs.kb, ke = zs.ke;
    if (kb >= ke) return;  // uniform per block
    const bool active = (i >= R) && (i < g.nnx - R) && (j >= R) && (j < g.nny - R);
    const int ic = clampi(i, 0, g.nnx - 1), jc = clampi(j, 0, g.nny - 1);
    const int col = jc * g.px + ic;
    const unsigned o_c = 4u * (unsigned)col;
    // x-halo: lanes tx<R load the left columns, lanes tx>=TX-R the right ones
    const bool xh = (tx < R) || (tx >= TX - R);
    const int lxh = tx < R ? tx : tx + 2 * R;
    const unsigned o_x =
        4u * (unsigned)(jc * g.px + (tx < R ? clampi(i0 - R + tx, 0, g.nnx - 1) : clampi(i0 + tx + R, 0, g.nnx - 1)));
    // y-halo: 2R rows above/below the tile, spread over the TY thread rows
    unsigned o_y[NYH];
    int l_y[NYH];
#pragma unroll
    for (int t = 0; t < NYH; ++t) {
        const int r = ty + t * TY;
        const int jj = r < R ? j0 - R + r : j0 + TY + (r - R);
        o_y[t] = 4u * (unsigned)(clampi(jj, 0, g.nny - 1) * g.px + ic);
        l_y[t] = r < R ? r : TY + r;
    }
    const bool is_src = active && col == src_col;
    // C-PML: x/y slab positions are fixed per column
    const int sxi = (AX & 1) && (tm & 1) && active ? slab_of(i, g.nnx, P.lo[0], P.hi[0]) : -1;
    const int syi = (AX & 2) && (tm & 2) && active ? slab_of(j, g.nny, P.lo[1], P.hi[1]) : -1;

    float c[R + 1];
#pragma unroll
    for (int m = 1; m <= R; ++m) c[m] = cf.c[m];
    const float hxy = active ? g.hx[ic] + g.hy[jc] : 0.f;
    // C-PML coefficients of the tile in LDS (x by column, y by row, z by plane,
    // the plane's entry written by one lane each step), so the update reads
    // them without a dependent global load
    __shared__ float4 pc_x[AX ? TX : 1], pc_y[AX ? TY : 1], pc_z[1];
    if constexpr (AX != 0) {
        const float4* cf4 = reinterpret_cast<const float4*>(P.psi + psi_total(P));
        if (ty == 0) pc_x[tx] = cf4[ic];
        if (ty == 1 && tx < TY) pc_y[tx] = cf4[g.nnx + clampi(j0 + tx, 0, g.nny - 1)];
    }

    // array bases at plane kb-R
    const long base = (long)(kb - R) * g.pxy;
    const long span = ((long)g.nnz * g.pxy - base) * 4;
    const unsigned pb = 4u * (unsigned)g.pxy;  // plane pitch in bytes
    using Acc = std::conditional_t<AX == 0, Rsrc, Base>;
    const Acc rvx = mkacc<Acc>(w.vx + base, span), rvy = mkacc<Acc>(w.vy + base, span), rvz = mkacc<Acc>(w.vz + base, span);
    const Acc rxx = mkacc<Acc>(w.txx + base, span), ryy = mkacc<Acc>(w.tyy + base, span), rzz = mkacc<Acc>(w.tzz + base, span);
    const Acc rxy = mkacc<Acc>(w.txy + base, span), rxz = mkacc<Acc>(w.txz + base, span), ryz = mkacc<Acc>(w.tyz + base, span);
    const Acc rtp = mkacc<Acc>(w.tp + base, span), rM = mkacc<Acc>(Mm + base, span), rU = mkacc<Acc>(MUm + base, span);
    // C-PML state (unused, and removed, when AX == 0)
    [[maybe_unused]] const Base rpsi = mkacc<Base>(P.psi, 0);

    float qx[2 * R + 1], qy[2 * R + 1], qz[2 * R + 1];
#pragma unroll
    for (int p = 0; p < 2 * R; ++p) {
        const unsigned so = (unsigned)p * pb;
        qx[p + 1] = bld(rvx, o_c, so);
        qy[p + 1] = bld(rvy, o_c, so);
        qz[p + 1] = bld(rvz, o_c, so);
    }
    // Complete the queue before the sweep: otherwise the compiler keeps these
    // loads pending into the loop and every step's queue shift waits on the
    // previous step's stores.
    __builtin_amdgcn_s_waitcnt(0);

    for (int k = kb; k < ke; ++k) {
#pragma unroll
        for (int p = 0; p < 2 * R; ++p) { qx[p] = qx[p + 1]; qy[p] = qy[p + 1]; qz[p] = qz[p + 1]; }
        const unsigned so = (unsigned)(k - kb + R) * pb;  // plane k
        const unsigned sf = so + (unsigned)R * pb;        // leading plane k+R
        qx[2 * R] = bld(rvx, o_c, sf);
        qy[2 * R] = bld(rvy, o_c, sf);
        qz[2 * R] = bld(rvz, o_c, sf);
        const float hzk = g.hz[k];  // issued with the plane's loads (a late load would stall the update)
        float hx0 = 0.f, hx1 = 0.f, hx2 = 0.f;
        if (xh) { hx0 = bld(rvx, o_x, so); hx1 = bld(rvy, o_x, so); hx2 = bld(rvz, o_x, so); }
        float hy0[NYH], hy1[NYH], hy2[NYH];
#pragma unroll
        for (int t = 0; t < NYH; ++t) {
            if (ty + t * TY < 2 * R) {
                hy0[t] = bld(rvx, o_y[t], so); hy1[t] = bld(rvy, o_y[t], so); hy2[t] = bld(rvz, o_y[t], so);
            }
        }
        // own stress state (issued before the barriers to overlap latency)
        float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f, stp = 0.f, M = 0.f, mu = 0.f;
        if (active) {
            sxx = bldo(rxx, o_c, so); syy = bldo(ryy, o_c, so); szz = bldo(rzz, o_c, so);
            sxy = bldo(rxy, o_c, so); sxz = bldo(rxz, o_c, so); syz = bldo(ryz, o_c, so);
            stp = bldo(rtp, o_c, so);
            M = bldo(rM, o_c, so); mu = bldo(rU, o_c, so);
        }
        // C-PML memory variables of this cell, prefetched likewise
        float ps[9];
        unsigned pqx = 0, pqy = 0, pqz = 0;  // slab offsets (< 2^32 floats)
        int szi = -1;
        // (every lane loads, at a clamped slab cell when outside the slab, so
        // no branch splits the loads; such lanes use zero memory variables,
        // which the zero coefficients there keep at zero, and store nothing)
        if constexpr (AX != 0) {
            szi = (AX & 4) && (tm & 4) && active ? slab_of(k, g.nnz, P.lo[2], P.hi[2]) : -1;
            if constexpr ((AX & 1) != 0) {
                if (tm & 1) {
                    pqx = ((unsigned)k * g.nny + jc) * P.w[0] + (unsigned)max(sxi, 0);
                    bldn<3>(rpsi, 12u * pqx, 4u * psi_stress(P, 0), ps);
                }
            }
            if constexpr ((AX & 2) != 0) {
                if (tm & 2) {
                    pqy = ((unsigned)k * P.w[1] + (unsigned)max(syi, 0)) * g.nnx + ic;
                    bldn<3>(rpsi, 12u * pqy, 4u * psi_stress(P, 1), ps + 3);
                }
            }
            if constexpr ((AX & 4) != 0) {
                if (tm & 4) {
                    pqz = ((unsigned)max(szi, 0) * g.nny + jc) * g.nnx + ic;
                    bldn<3>(rpsi, 12u * pqz, 4u * psi_stress(P, 2), ps + 6);
                }
            }
        }
        [[maybe_unused]] float4 zq{};
        if constexpr ((AX & 4) != 0) {
            if ((tm & 4) && tx == 0 && ty == 0) zq = reinterpret_cast<const float4*>(P.psi + psi_total(P) + 4u * (g.nnx + g.nny))[k];
        }
        __syncthreads();  // previous plane's LDS reads are done
        if constexpr ((AX & 4) != 0) {
            if ((tm & 4) && tx == 0 && ty == 0) pc_z[0] = zq;
        }
        sx[ty + R][tx + R] = qx[R];
        sy[ty + R][tx + R] = qy[R];
        sz[ty + R][tx + R] = qz[R];
        if (xh) { sx[ty + R][lxh] = hx0; sy[ty + R][lxh] = hx1; sz[ty + R][lxh] = hx2; }
#pragma unroll
        for (int t = 0; t < NYH; ++t) {
            if (ty + t * TY < 2 * R) {
                sx[l_y[t]][tx + R] = hy0[t]; sy[l_y[t]][tx + R] = hy1[t]; sz[l_y[t]][tx + R] = hy2[t];
            }
        }
        __syncthreads();
        // Every lane computes (inactive lanes on clamped, in-range data) and
        // only the stores are guarded: a load consumed on some control paths
        // only would make the compiler wait for this plane's stores before
        // the next plane's queue shift, serialising two memory round trips.
        {
            const int ci = tx + R, cj = ty + R;
            float dxvx = 0.f, dyvy = 0.f, dzvz = 0.f, dyvx = 0.f, dxvy = 0.f, dzvx = 0.f, dxvz = 0.f,
                  dzvy = 0.f, dyvz = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m) {
                dxvx += c[m] * (sx[cj][ci + m] - sx[cj][ci - m + 1]);          // D+x vx
                dyvy += c[m] * (sy[cj + m][ci] - sy[cj - m + 1][ci]);          // D+y vy
                dzvz += c[m] * (qz[R + m - 1] - qz[R - m]);                    // D-z vz
                dyvx += c[m] * (sx[cj + m - 1][ci] - sx[cj - m][ci]);          // D-y vx
                dxvy += c[m] * (sy[cj][ci + m - 1] - sy[cj][ci - m]);          // D-x vy
                dzvx += c[m] * (qx[R + m] - qx[R - m + 1]);                    // D+z vx
                dxvz += c[m] * (sz[cj][ci + m - 1] - sz[cj][ci - m]);          // D-x vz
                dzvy += c[m] * (qy[R + m] - qy[R - m + 1]);                    // D+z vy
                dyvz += c[m] * (sz[cj + m - 1][ci] - sz[cj - m][ci]);          // D-y vz
            }
            if (is_src && k == src_k && it < wl) {
                const float s = wavelet[it];
                sxx += s; syy += s; szz += s;
            }
            float exx = g.cx * dxvx, eyy = g.cy * dyvy, ezz = g.cz * dzvz;
            float e_yx = g.cy * dyvx, e_xy = g.cx * dxvy, e_zx = g.cz * dzvx, e_xz = g.cx * dxvz;
            float e_zy = g.cz * dzvy, e_yz = g.cy * dyvz;
            if constexpr (AX != 0) {
                if constexpr ((AX & 1) != 0) if (tm & 1) {
                    const float4 cq = pc_x[tx];  // (a_full, b_full, a_half, b_half) of x
                    for (int q = 0; q < 3; ++q) ps[q] = sxi >= 0 ? ps[q] : 0.f;
                    exx = pml_upd(cq.x, cq.y, exx, ps[0]);
                    e_xy = pml_upd(cq.z, cq.w, e_xy, ps[1]);
                    e_xz = pml_upd(cq.z, cq.w, e_xz, ps[2]);
                }
                if constexpr ((AX & 2) != 0) if (tm & 2) {
                    const float4 cq = pc_y[ty];
                    for (int q = 3; q < 6; ++q) ps[q] = syi >= 0 ? ps[q] : 0.f;
                    eyy = pml_upd(cq.x, cq.y, eyy, ps[3]);
                    e_yx = pml_upd(cq.z, cq.w, e_yx, ps[4]);
                    e_yz = pml_upd(cq.z, cq.w, e_yz, ps[5]);
                }
                if constexpr ((AX & 4) != 0) if (tm & 4) {
                    const float4 cq = pc_z[0];
                    for (int q = 6; q < 9; ++q) ps[q] = szi >= 0 ? ps[q] : 0.f;
                    ezz = pml_upd(cq.x, cq.y, ezz, ps[6]);
                    e_zx = pml_upd(cq.z, cq.w, e_zx, ps[7]);
                    e_zy = pml_upd(cq.z, cq.w, e_zy, ps[8]);
                }
            }
            const float h = hxy + hzk;
            const float a = 1.0f / (1.0f + h), b = 1.0f - h;
            const float lam = M - 2.0f * mu;
            const float nxx = a * (b * sxx + M * exx + lam * (eyy + ezz));
            const float nyy = a * (b * syy + M * eyy + lam * (exx + ezz));
            const float nzz = a * (b * szz + M * ezz + lam * (exx + eyy));
            const float ntp = a * (b * stp + M * (exx + eyy + ezz));
            const float nxy = a * (b * sxy + mu * (e_yx + e_xy));
            const float nxz = a * (b * sxz + mu * (e_zx + e_xz));
            const float nyz = a * (b * syz + mu * (e_zy + e_yz));
            if (active) {
                bst(rxx, o_c, so, nxx); bst(ryy, o_c, so, nyy); bst(rzz, o_c, so, nzz); bst(rtp, o_c, so, ntp);
                bst(rxy, o_c, so, nxy); bst(rxz, o_c, so, nxz); bst(ryz, o_c, so, nyz);
            }
            if constexpr ((AX & 1) != 0) if (sxi >= 0) bstn<3>(rpsi, 12u * pqx, 4u * psi_stress(P, 0), ps);
            if constexpr ((AX & 2) != 0) if (syi >= 0) bstn<3>(rpsi, 12u * pqy, 4u * psi_stress(P, 1), ps + 3);
            if constexpr ((AX & 4) != 0) if (szi >= 0) bstn<3>(rpsi, 12u * pqz, 4u * psi_stress(P, 2), ps + 6);
        }
    }
}

// AX: C-PML axis group of the tiles this launch covers (bit 0 x, 1 y, 2 z;
// 0: no memory variables).  Radius-4 occupancy targets: 4 waves per SIMD for
// the single-axis groups (they fit 128 VGPRs without spilling), 3 for the
// multi-axis groups.
template <int R, int TX, int TY, int AX>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(R != 4 || AX == 0 ? 1 : (AX == 1 || AX == 2 || AX == 4) ? 4 : 3)))
void k_stress(Wave w, const float* __restrict__ Mm,
                                                   const float* __restrict__ MUm, Geo g, Coef cf,
                                                   int src_k, int src_col, const float* __restrict__ wavelet,
                                                   int wl, int it, Pml P, const int* __restrict__ blist) {
    // blist (optional): tile ids of this launch (region-split launches; bits
    // 28..30 the C-PML axes whose slabs the tile touches); else one tile per workgroup
    const int e = blist ? blist[blockIdx.x] : (int)blockIdx.x;
    const int blk = blist ? (e & kTileIdMask) : e, tm = blist ? (e >> 28) & 7 : 7;
    stress_tile<R, TX, TY, AX>(w, Mm, MUm, g, cf, src_k, src_col, wavelet, wl, it, P, blk, tm);
}

// ---------------------------------------------------------------------------
// k_velocity: fwd_vx_3D + fwd_vy_3D + fwd_vz_3D (:100-104) + fwd_vxp_vzp_3D
// (:118), with vector S = v - vP left implicit (computed where it is used),
// fused with
//   MODE_EXCITE: cuda_cal_excitation_amp_time_new (:282): keep the step of
//                maximum |vP| and tp, vP at that step (2D form:
//                elastic_2D_kernel_3.cu:30-48, strict '<': first maximum wins);
//   MODE_IMAGE : imaging_correlation_ex + 2x imaging_inner_product_ex (:442-446):
//                where ex_time == it, I_tp = tp/ex_tp, I_PP = ex_vP.vP/|ex_vP|^2,
//                I_PS = ex_vP.vS/|ex_vP|^2 (elastic_2D_kernel_3.cu:542-580);
//                a zero denominator images 0.
// HBM per cell-update: reads stress+tp(7), rw v+vP(12), reads 1/rho(1) = 80 B,
// +4 B ex_amp (EXCITE) or ex_time (IMAGE).
// ---------------------------------------------------------------------------
template <int R, int TX, int TY, int MODE, int AX>
__device__ __forceinline__ void velocity_tile(const Wave& w, const float* __restrict__ Bm, const Geo& g,
                                              const Coef& cf, const Excite& ex, int it, const Pml& P, int blk, int tm) {
    constexpr int LX = TX + 2 * R, LY = TY + 2 * R;
    constexpr int NYH = (2 * R + TY - 1) / TY;
    __shared__ float s_xx[TY][LX];   // x-halo only
    __shared__ float s_yy[LY][TX];   // y-halo only
    __shared__ float s_xy[LY][LX];   // x and y halo
    __shared__ float s_xz[TY][LX];   // x-halo only
    __shared__ float s_yz[LY][TX];   // y-halo only
    __shared__ float s_tp[LY][LX];   // x and y halo

    const TileId tid = tile_of_block(g, blk);
    if (!tid.valid) return;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = g.x0 + tid.bx * TX, j0 = tid.by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const ZSpan zs = chunk_span(g, R, tid.bz);
    const int kb = zs.kb, ke = zs.ke;
    if (kb >= ke) return;
    const bool active = (i >= R) && (i < g.nnx - R) && (j >= R) && (j < g.nny - R);
    const int ic = clampi(i, 0, g.nnx - 1), jc = clampi(j, 0, g.nny - 1);
    const unsigned o_c = 4u * (unsigned)(jc * g.px + ic);
    const bool xh = (tx < R) || (tx >= TX - R);
    const int lxh = tx < R ? tx : tx + 2 * R;
    const unsigned o_x =
        4u * (unsigned)(jc * g.px + (tx < R ? clampi(i0 - R + tx, 0, g.nnx - 1) : clampi(i0 + tx + R, 0, g.nnx - 1)));
    unsigned o_y[NYH];
    int l_y[NYH];
#pragma unroll
    for (int t = 0; t < NYH; ++t) {
        const int r = ty + t * TY;
        const int jj = r < R ? j0 - R + r : j0 + TY + (r - R);
        o_y[t] = 4u * (unsigned)(clampi(jj, 0, g.nny - 1) * g.px + ic);
        l_y[t] = r < R ? r : TY + r;
    }

    float c[R + 1];
#pragma unroll
    for (int m = 1; m <= R; ++m) c[m] = cf.c[m];
    const float hxy = active ? g.hx[ic] + g.hy[jc] : 0.f;
    // C-PML coefficients in LDS, as in the stress tile
    __shared__ float4 pc_x[AX ? TX : 1], pc_y[AX ? TY : 1], pc_z[1];
    if constexpr (AX != 0) {
        const float4* cf4 = reinterpret_cast<const float4*>(P.psi + psi_total(P));
        if (ty == 0) pc_x[tx] = cf4[ic];
        if (ty == 1 && tx < TY) pc_y[tx] = cf4[g.nnx + clampi(j0 + tx, 0, g.nny - 1)];
    }
    const int sxi = (AX & 1) && (tm & 1) && active ? slab_of(i, g.nnx, P.lo[0], P.hi[0]) : -1;
    const int syi = (AX & 2) && (tm & 2) && active ? slab_of(j, g.nny, P.lo[1], P.hi[1]) : -1;

    const long base = (long)(kb - R) * g.pxy;
    const long span = ((long)g.nnz * g.pxy - base) * 4;
    const unsigned pb = 4u * (unsigned)g.pxy;
    using Acc = std::conditional_t<AX == 0, Rsrc, Base>;
    const Acc rvx = mkacc<Acc>(w.vx + base, span), rvy = mkacc<Acc>(w.vy + base, span), rvz = mkacc<Acc>(w.vz + base, span);
    const Acc rpx = mkacc<Acc>(w.vxp + base, span), rpy = mkacc<Acc>(w.vyp + base, span), rpz = mkacc<Acc>(w.vzp + base, span);
    const Acc rxx = mkacc<Acc>(w.txx + base, span), ryy = mkacc<Acc>(w.tyy + base, span), rzz = mkacc<Acc>(w.tzz + base, span);
    const Acc rxy = mkacc<Acc>(w.txy + base, span), rxz = mkacc<Acc>(w.txz + base, span), ryz = mkacc<Acc>(w.tyz + base, span);
    const Acc rtp = mkacc<Acc>(w.tp + base, span), rB = mkacc<Acc>(Bm + base, span);
    [[maybe_unused]] const Base rpsi = mkacc<Base>(P.psi, 0);

    float qxz[2 * R + 1], qyz[2 * R + 1], qzz[2 * R + 1], qtp[2 * R + 1];
#pragma unroll
    for (int p = 0; p < 2 * R; ++p) {
        const unsigned so = (unsigned)p * pb;
        qxz[p + 1] = bld(rxz, o_c, so);
        qyz[p + 1] = bld(ryz, o_c, so);
        qzz[p + 1] = bld(rzz, o_c, so);
        qtp[p + 1] = bld(rtp, o_c, so);
    }
    __builtin_amdgcn_s_waitcnt(0);  // see the stress tile

    for (int k = kb; k < ke; ++k) {
#pragma unroll
        for (int p = 0; p < 2 * R; ++p) {
            qxz[p] = qxz[p + 1]; qyz[p] = qyz[p + 1]; qzz[p] = qzz[p + 1]; qtp[p] = qtp[p + 1];
        }
        const unsigned so = (unsigned)(k - kb + R) * pb;
        const unsigned sf = so + (unsigned)R * pb;
        qxz[2 * R] = bld(rxz, o_c, sf);
        qyz[2 * R] = bld(ryz, o_c, sf);
        qzz[2 * R] = bld(rzz, o_c, sf);
        qtp[2 * R] = bld(rtp, o_c, sf);
        const float hzk = g.hz[k];
        const float cxx = bld(rxx, o_c, so), cyy = bld(ryy, o_c, so), cxy = bld(rxy, o_c, so);
        float hxx = 0.f, hxy_ = 0.f, hxz = 0.f, htp = 0.f;
        if (xh) {
            hxx = bld(rxx, o_x, so); hxy_ = bld(rxy, o_x, so); hxz = bld(rxz, o_x, so); htp = bld(rtp, o_x, so);
        }
        float yh_yy[NYH], yh_xy[NYH], yh_yz[NYH], yh_tp[NYH];  // y-halo rows of this plane
#pragma unroll
        for (int t = 0; t < NYH; ++t) {
            if (ty + t * TY < 2 * R) {
                yh_yy[t] = bld(ryy, o_y[t], so); yh_xy[t] = bld(rxy, o_y[t], so);
                yh_yz[t] = bld(ryz, o_y[t], so); yh_tp[t] = bld(rtp, o_y[t], so);
            }
        }
        float vx = 0.f, vy = 0.f, vz = 0.f, vxp = 0.f, vyp = 0.f, vzp = 0.f, bo = 0.f;
        if (active) {
            vx = bldo(rvx, o_c, so); vy = bldo(rvy, o_c, so); vz = bldo(rvz, o_c, so);
            vxp = bldo(rpx, o_c, so); vyp = bldo(rpy, o_c, so); vzp = bldo(rpz, o_c, so);
            bo = bldo(rB, o_c, so);
        }
        // the cell's excitation amplitude (source pass) or time (receiver pass),
        // prefetched with the plane's loads: read after the stores it would wait for them
        [[maybe_unused]] float ex_amp_c = 0.f;
        [[maybe_unused]] int ex_time_c = -1;
        if constexpr (MODE == MODE_EXCITE) ex_amp_c = bldo(mkacc<Acc>(ex_array(ex, EX_AMP) + base, span), o_c, so);
        if constexpr (MODE == MODE_IMAGE) ex_time_c = bldi(mkacc<Acc>(ex_array(ex, EX_TIME) + base, span), o_c, so);
        float ps[12];
        unsigned pqx = 0, pqy = 0, pqz = 0;  // slab offsets (< 2^32 floats)
        int szi = -1;
        if constexpr (AX != 0) {  // unconditional loads at clamped slab cells, as in the stress tile
            szi = (AX & 4) && (tm & 4) && active ? slab_of(k, g.nnz, P.lo[2], P.hi[2]) : -1;
            if constexpr ((AX & 1) != 0) {
                if (tm & 1) {
                    pqx = ((unsigned)k * g.nny + jc) * P.w[0] + (unsigned)max(sxi, 0);
                    bldn<4>(rpsi, 16u * pqx, 4u * psi_velocity(P, 0), ps);
                }
            }
            if constexpr ((AX & 2) != 0) {
                if (tm & 2) {
                    pqy = ((unsigned)k * P.w[1] + (unsigned)max(syi, 0)) * g.nnx + ic;
                    bldn<4>(rpsi, 16u * pqy, 4u * psi_velocity(P, 1), ps + 4);
                }
            }
            if constexpr ((AX & 4) != 0) {
                if (tm & 4) {
                    pqz = ((unsigned)max(szi, 0) * g.nny + jc) * g.nnx + ic;
                    bldn<4>(rpsi, 16u * pqz, 4u * psi_velocity(P, 2), ps + 8);
                }
            }
        }
        [[maybe_unused]] float4 zq{};
        if constexpr ((AX & 4) != 0) {
            if ((tm & 4) && tx == 0 && ty == 0) zq = reinterpret_cast<const float4*>(P.psi + psi_total(P) + 4u * (g.nnx + g.nny))[k];
        }
        __syncthreads();
        if constexpr ((AX & 4) != 0) {
            if ((tm & 4) && tx == 0 && ty == 0) pc_z[0] = zq;
        }
        s_xx[ty][tx + R] = cxx;
        s_yy[ty + R][tx] = cyy;
        s_xy[ty + R][tx + R] = cxy;
        s_xz[ty][tx + R] = qxz[R];
        s_yz[ty + R][tx] = qyz[R];
        s_tp[ty + R][tx + R] = qtp[R];
        if (xh) {
            s_xx[ty][lxh] = hxx; s_xy[ty + R][lxh] = hxy_; s_xz[ty][lxh] = hxz; s_tp[ty + R][lxh] = htp;
        }
#pragma unroll
        for (int t = 0; t < NYH; ++t) {
            if (ty + t * TY < 2 * R) {
                const int lr = l_y[t];
                s_yy[lr][tx] = yh_yy[t]; s_xy[lr][tx + R] = yh_xy[t]; s_yz[lr][tx] = yh_yz[t]; s_tp[lr][tx + R] = yh_tp[t];
            }
        }
        __syncthreads();
        {  // all lanes compute, only the stores are guarded (see the stress tile)
            const int ci = tx + R, cj = ty + R;
            float a1 = 0.f, a2 = 0.f, a3 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f,
                  p1 = 0.f, p2 = 0.f, p3 = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m) {
                a1 += c[m] * (s_xx[ty][ci + m - 1] - s_xx[ty][ci - m]);        // D-x sxx
                a2 += c[m] * (s_xy[cj + m][ci] - s_xy[cj - m + 1][ci]);        // D+y sxy
                a3 += c[m] * (qxz[R + m - 1] - qxz[R - m]);                    // D-z sxz
                b1 += c[m] * (s_xy[cj][ci + m] - s_xy[cj][ci - m + 1]);        // D+x sxy
                b2 += c[m] * (s_yy[cj + m - 1][tx] - s_yy[cj - m][tx]);        // D-y syy
                b3 += c[m] * (qyz[R + m - 1] - qyz[R - m]);                    // D-z syz
                d1 += c[m] * (s_xz[ty][ci + m] - s_xz[ty][ci - m + 1]);        // D+x sxz
                d2 += c[m] * (s_yz[cj + m][tx] - s_yz[cj - m + 1][tx]);        // D+y syz
                d3 += c[m] * (qzz[R + m] - qzz[R - m + 1]);                    // D+z szz
                p1 += c[m] * (s_tp[cj][ci + m - 1] - s_tp[cj][ci - m]);        // D-x tp
                p2 += c[m] * (s_tp[cj + m - 1][ci] - s_tp[cj - m][ci]);        // D-y tp
                p3 += c[m] * (qtp[R + m] - qtp[R - m + 1]);                    // D+z tp
            }
            float gxx = g.cx * a1, gyx = g.cy * a2, gzx = g.cz * a3;
            float gxy = g.cx * b1, gyy = g.cy * b2, gzy = g.cz * b3;
            float gxz = g.cx * d1, gyz = g.cy * d2, gzz = g.cz * d3;
            float px = g.cx * p1, py = g.cy * p2, pz = g.cz * p3;
            if constexpr (AX != 0) {
                if constexpr ((AX & 1) != 0) if (tm & 1) {
                    const float4 cq = pc_x[tx];  // (a_full, b_full, a_half, b_half) of x
                    for (int q = 0; q < 4; ++q) ps[q] = sxi >= 0 ? ps[q] : 0.f;
                    gxx = pml_upd(cq.z, cq.w, gxx, ps[0]);
                    gxy = pml_upd(cq.x, cq.y, gxy, ps[1]);
                    gxz = pml_upd(cq.x, cq.y, gxz, ps[2]);
                    px = pml_upd(cq.z, cq.w, px, ps[3]);
                }
                if constexpr ((AX & 2) != 0) if (tm & 2) {
                    const float4 cq = pc_y[ty];
                    for (int q = 4; q < 8; ++q) ps[q] = syi >= 0 ? ps[q] : 0.f;
                    gyx = pml_upd(cq.x, cq.y, gyx, ps[4]);
                    gyy = pml_upd(cq.z, cq.w, gyy, ps[5]);
                    gyz = pml_upd(cq.x, cq.y, gyz, ps[6]);
                    py = pml_upd(cq.z, cq.w, py, ps[7]);
                }
                if constexpr ((AX & 4) != 0) if (tm & 4) {
                    const float4 cq = pc_z[0];
                    for (int q = 8; q < 12; ++q) ps[q] = szi >= 0 ? ps[q] : 0.f;
                    gzx = pml_upd(cq.x, cq.y, gzx, ps[8]);
                    gzy = pml_upd(cq.x, cq.y, gzy, ps[9]);
                    gzz = pml_upd(cq.z, cq.w, gzz, ps[10]);
                    pz = pml_upd(cq.z, cq.w, pz, ps[11]);
                }
            }
            const float h = hxy + hzk;
            const float a = 1.0f / (1.0f + h), b = 1.0f - h;
            const float nvx = a * (b * vx + bo * (gxx + gyx + gzx));
            const float nvy = a * (b * vy + bo * (gxy + gyy + gzy));
            const float nvz = a * (b * vz + bo * (gxz + gyz + gzz));
            const float nxp = a * (b * vxp + bo * px);
            const float nyp = a * (b * vyp + bo * py);
            const float nzp = a * (b * vzp + bo * pz);
            if constexpr ((AX & 1) != 0) if (sxi >= 0) bstn<4>(rpsi, 16u * pqx, 4u * psi_velocity(P, 0), ps);
            if constexpr ((AX & 2) != 0) if (syi >= 0) bstn<4>(rpsi, 16u * pqy, 4u * psi_velocity(P, 1), ps + 4);
            if constexpr ((AX & 4) != 0) if (szi >= 0) bstn<4>(rpsi, 16u * pqz, 4u * psi_velocity(P, 2), ps + 8);
            if (active) {
            bst(rvx, o_c, so, nvx); bst(rvy, o_c, so, nvy); bst(rvz, o_c, so, nvz);
            bst(rpx, o_c, so, nxp); bst(rpy, o_c, so, nyp); bst(rpz, o_c, so, nzp);
            // excitation / imaging state: the per-cell test reads one word
            // through a buffer; the rare conditional accesses use plain pointers
            const long n = base + (long)(so / 4u) + (long)(o_c / 4u);
            if constexpr (MODE == MODE_EXCITE) {
                const float amp = sqrtf(nxp * nxp + nyp * nyp + nzp * nzp);
                if (ex_amp_c < amp) {
                    reinterpret_cast<int*>(ex_array(ex, EX_TIME))[n] = it;
                    ex_array(ex, EX_AMP)[n] = amp; ex_array(ex, EX_TP)[n] = qtp[R];
                    ex_array(ex, EX_VXP)[n] = nxp; ex_array(ex, EX_VYP)[n] = nyp; ex_array(ex, EX_VZP)[n] = nzp;
                }
            } else if constexpr (MODE == MODE_IMAGE) {
                if (ex_time_c == it) {
                    const float etp = ex_array(ex, EX_TP)[n];
                    ex_array(ex, EX_IMG_TP)[n] = etp != 0.f ? qtp[R] / etp : 0.f;
                    const float ex_ = ex_array(ex, EX_VXP)[n], ey_ = ex_array(ex, EX_VYP)[n], ez_ = ex_array(ex, EX_VZP)[n];
                    const float den = ex_ * ex_ + ey_ * ey_ + ez_ * ez_;
                    const float pp = ex_ * nxp + ey_ * nyp + ez_ * nzp;
                    const float ips = ex_ * (nvx - nxp) + ey_ * (nvy - nyp) + ez_ * (nvz - nzp);
                    ex_array(ex, EX_IMG_PP)[n] = den > 0.f ? pp / den : 0.f;
                    ex_array(ex, EX_IMG_PS)[n] = den > 0.f ? ips / den : 0.f;
                }
            }
            }
        }
    }
}

// Occupancy target of the radius-4 C-PML velocity variants: 3 waves per SIMD
// (<= 168 VGPRs; the prefetched memory variables would push the all-axes
// group just past it, and 4 waves would spill).  Radius 6 keeps its natural
// occupancy (forcing it would spill).
template <int R, int TX, int TY, int MODE, int AX>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(R != 4 || AX == 0 ? 1 : 3)))
void k_velocity(Wave w, const float* __restrict__ Bm, Geo g,
                                                     Coef cf, Excite ex, int it, Pml P,
                                                     const int* __restrict__ blist) {
    const int e = blist ? blist[blockIdx.x] : (int)blockIdx.x;
    const int blk = blist ? (e & kTileIdMask) : e, tm = blist ? (e >> 28) & 7 : 7;
    velocity_tile<R, TX, TY, MODE, AX>(w, Bm, g, cf, ex, it, P, blk, tm);
}

// ---------------------------------------------------------------------------
// Batched launches: many shots (each its own window grid, of any size) in one
// launch per kernel.  table[b] = (shot, block of that shot's tiling, with the
// C-PML axis mask in bits 28..30 as in the one-shot tile lists); the
// grid either covers the table once or is persistent (fewer workgroups, each
// striding over the table).  The time step is read from device memory so that
// one captured step can be replayed (k_step_advance moves it).
// ---------------------------------------------------------------------------
struct ShotStep {
    Wave w;
    Excite ex;
    Pml P;
    Geo g;
    const float* M;
    const float* MU;
    const float* B;
    int src_k, src_col;
};

// LOOP: persistent grid (workgroups stride over the table); without it the
// grid covers the table once, one entry per workgroup, which frees the loop's
// scalar registers (the plain velocity kernels then keep their occupancy).
template <int R, int TX, int TY, int AX, bool LOOP>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(R != 4 || AX == 0 ? 1 : (AX == 1 || AX == 2 || AX == 4) ? 4 : 3)))
void k_stress_batch(const ShotStep* __restrict__ shots,
                                                         const int2* __restrict__ table, int total, Coef cf,
                                                         const float* __restrict__ wavelet, int wl,
                                                         const int* __restrict__ itp, int source_on) {
    const int it = *itp;
    for (int b = (int)blockIdx.x; b < total; b += (int)gridDim.x) {
        const int2 t = table[b];
        const ShotStep& s = shots[t.x];
        stress_tile<R, TX, TY, AX>(s.w, s.M, s.MU, s.g, cf, source_on ? s.src_k : -1, source_on ? s.src_col : -1,
                                     wavelet, wl, it, s.P, t.y & kTileIdMask, (t.y >> 28) & 7);
        if constexpr (!LOOP) break;
    }
}

template <int R, int TX, int TY, int MODE, int AX, bool LOOP>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(R != 4 || AX == 0 ? 1 : 3)))
void k_velocity_batch(const ShotStep* __restrict__ shots,
                                                           const int2* __restrict__ table, int total, Coef cf,
                                                           const int* __restrict__ itp) {
    const int it = *itp;
    for (int b = (int)blockIdx.x; b < total; b += (int)gridDim.x) {
        const int2 t = table[b];
        const ShotStep& s = shots[t.x];
        velocity_tile<R, TX, TY, MODE, AX>(s.w, s.B, s.g, cf, s.ex, it, s.P, t.y & kTileIdMask, (t.y >> 28) & 7);
        if constexpr (!LOOP) break;
    }
}

// Receiver Poynting vectors at each cell's imaging step (erm_params.poynting),
// the body of cuda_cal_receiver_poyn_3D, whose call is commented out at
// 3D_forward_or_back_together.cu:432 and whose body is absent; the 3D form of
// sum_poynting (2D_elastic_RTM/elastic_2D_kernel_2.cu:8-51):
//   rp = -tp vP,   rs = -(sigma - tp I) vS,   vS = v - vP,
// every product at the cell's own index, as sum_poynting.  Runs after the
// velocity update of receiver step it, so the fields are those the imaging
// condition of the same step reads.  Pointwise over the n allocated cells of
// a volume; ex_time is -1 outside the updated cells.
[[maybe_unused]] static __global__ void k_poynting(Wave w, const int* __restrict__ ex_time, float* __restrict__ rpx,
                                                   float* __restrict__ rpy, float* __restrict__ rpz,
                                                   float* __restrict__ rsx, float* __restrict__ rsy,
                                                   float* __restrict__ rsz, long n, int it) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        if (ex_time[q] != it) continue;
        const float tp = w.tp[q], px = w.vxp[q], py = w.vyp[q], pz = w.vzp[q];
        const float sx = w.vx[q] - px, sy = w.vy[q] - py, sz = w.vz[q] - pz;
        const float xx = w.txx[q] - tp, yy = w.tyy[q] - tp, zz = w.tzz[q] - tp;
        const float xy = w.txy[q], xz = w.txz[q], yz = w.tyz[q];
        rpx[q] = -tp * px; rpy[q] = -tp * py; rpz[q] = -tp * pz;
        rsx[q] = -(xx * sx + xy * sy + xz * sz);
        rsy[q] = -(xy * sx + yy * sy + yz * sz);
        rsz[q] = -(xz * sx + yz * sy + zz * sz);
    }
}

// Receivers of all shots: shot_of[r] names the shot of global receiver r,
// RecShot holds that shot's fields, record and the first global index.
struct RecShot {
    float *vx, *vy, *vz;
    const long* idx;
    float* rec;   // [3][lt][nr]
    int nr, first;
    long comp;    // lt * nr
};
[[maybe_unused]] static __global__ void k_receivers_batch(const RecShot* __restrict__ shots, const int* __restrict__ shot_of, int total,
                                  const int* __restrict__ itp, int inject) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= total) return;
    const RecShot& s = shots[shot_of[r]];
    const int q = r - s.first;
    const long n = s.idx[q], off = (long)(*itp) * s.nr + q;
    if (inject) {
        atomicAdd(&s.vx[n], s.rec[off]);
        atomicAdd(&s.vy[n], s.rec[s.comp + off]);
        atomicAdd(&s.vz[n], s.rec[2 * s.comp + off]);
    } else {
        s.rec[off] = s.vx[n];
        s.rec[s.comp + off] = s.vy[n];
        s.rec[2 * s.comp + off] = s.vz[n];
    }
}

// Boundary strips for source-wavefield reconstruction: the R-cell face slabs
// just outside the model region, where the stencils of model-region cells
// reach.  Gather (save) or scatter (restore) nf fields at the listed padded
// indices; buf layout [nf][n].
struct StripFields {
    float* f[7];
};
[[maybe_unused]] static __global__ void k_strip(StripFields F, int nf, const int* __restrict__ idx, int n,
                                                float* __restrict__ buf, int restore) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int q = idx[t];
    for (int a = 0; a < nf; ++a) {
        if (restore) F.f[a][q] = buf[(long)a * n + t];
        else buf[(long)a * n + t] = F.f[a][q];
    }
}

// Zero-lag cross-correlation imaging of a source and a receiver wavefield over
// the model region (origin (ox, oy, oz), size nx x ny x nz of the padded grid):
//   I_tp += tp_s tp_r,  I_PP += vP_s . vP_r,  I_PS += vP_s . (v_r - vP_r).
[[maybe_unused]] static __global__ void k_xcorr(Wave s, Wave r, float* __restrict__ itp, float* __restrict__ ipp,
                                                float* __restrict__ ips, int nx, int ny, int nz, int ox, int oy,
                                                int oz, int px, long pxy) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)nx * ny * nz) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((long)nx * ny));
    const long n = (long)(k + oz) * pxy + (long)(j + oy) * px + (i + ox);
    const float sx = s.vxp[n], sy = s.vyp[n], sz = s.vzp[n];
    const float rx = r.vxp[n], ry = r.vyp[n], rz = r.vzp[n];
    itp[n] += s.tp[n] * r.tp[n];
    ipp[n] += sx * rx + sy * ry + sz * rz;
    ips[n] += sx * (r.vx[n] - rx) + sy * (r.vy[n] - ry) + sz * (r.vz[n] - rz);
}

// Decimated source history (erm_forward_smoothed_history): tp, vxp, vyp, vzp
// of the model region (i fastest) after a source step -> h [4][nx*ny*nz].
[[maybe_unused]] static __global__ void k_hist_save(Wave s, float* __restrict__ h, int nx, int ny, int nz, int ox,
                                                    int oy, int oz, int px, long pxy) {
    const long nm = (long)nx * ny * nz;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nm; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((long)nx * ny));
        const long n = (long)(k + oz) * pxy + (long)(j + oy) * px + (i + ox);
        h[t] = s.tp[n]; h[nm + t] = s.vxp[n]; h[2 * nm + t] = s.vyp[n]; h[3 * nm + t] = s.vzp[n];
    }
}
// Zero-lag cross-correlation with a stored source step (k_xcorr with the
// source read from h), each product weighted by w (the decimation factor).
[[maybe_unused]] static __global__ void k_xcorr_hist(const float* __restrict__ h, Wave r, float* __restrict__ itp,
                                                     float* __restrict__ ipp, float* __restrict__ ips, int nx, int ny,
                                                     int nz, int ox, int oy, int oz, int px, long pxy, float w) {
    const long nm = (long)nx * ny * nz;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nm; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((long)nx * ny));
        const long n = (long)(k + oz) * pxy + (long)(j + oy) * px + (i + ox);
        const float st = h[t], sx = h[nm + t], sy = h[2 * nm + t], sz = h[3 * nm + t];
        const float rx = r.vxp[n], ry = r.vyp[n], rz = r.vzp[n];
        itp[n] += w * (st * r.tp[n]);
        ipp[n] += w * (sx * rx + sy * ry + sz * rz);
        ips[n] += w * (sx * (r.vx[n] - rx) + sy * (r.vy[n] - ry) + sz * (r.vz[n] - rz));
    }
}

// it += delta (one lane; a plain vector store).
[[maybe_unused]] static __global__ void k_step_advance(int* itp, int delta) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *itp += delta;
}

// Stack a window's model-region image into the full-model image (x fastest,
// dense [nz][ny][nx]) at lateral offset (x0, y0); img is the window's padded
// volume with model origin (ox, oy, oz) and pitches px, pxy.
[[maybe_unused]] static __global__ void k_stack(const float* __restrict__ img, int wx, int wy, int nz, int ox, int oy, int oz, int px,
                        long pxy, float* __restrict__ out, int nx, int ny, int x0, int y0) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)wx * wy * nz) return;
    const int i = (int)(t % wx), j = (int)((t / wx) % wy), k = (int)(t / ((long)wx * wy));
    out[((long)k * ny + (y0 + j)) * nx + (x0 + i)] += img[(long)(k + oz) * pxy + (long)(j + oy) * px + (i + ox)];
}

// ---------------------------------------------------------------------------
// Receivers: write_or_add_shot_3D_surface_three (:127 record, :328 inject).
// rec layout [3][lt][nr]; idx[r] is the padded-grid index of receiver r.
// ---------------------------------------------------------------------------
// direct-wave mute of a record [3][lt][nr]: samples it < cut[r] become 0
[[maybe_unused]] static __global__ void k_mute(float* __restrict__ rec, int nr, long lt, const int* __restrict__ cut) {
    const long n = 3 * lt * nr;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int r = (int)(t % nr);
        const long it = (t / nr) % lt;
        if (it < cut[r]) rec[t] = 0.f;
    }
}
[[maybe_unused]] static __global__ void k_record(const float* __restrict__ vx, const float* __restrict__ vy,
                         const float* __restrict__ vz, const long* __restrict__ idx, int nr,
                         float* __restrict__ rec, long comp_stride, long it_off) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const long n = idx[r];
    rec[it_off + r] = vx[n];
    rec[comp_stride + it_off + r] = vy[n];
    rec[2 * comp_stride + it_off + r] = vz[n];
}

[[maybe_unused]] static __global__ void k_inject(float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz,
                         const long* __restrict__ idx, int nr, const float* __restrict__ rec,
                         long comp_stride, long it_off) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const long n = idx[r];
    atomicAdd(&vx[n], rec[it_off + r]);
    atomicAdd(&vy[n], rec[comp_stride + it_off + r]);
    atomicAdd(&vz[n], rec[2 * comp_stride + it_off + r]);
}

// P/S decomposition by divergence and curl of v, the 3D form of decom
// (elastic_2D_kernel_2.cu:563-669) with the staggered positions above:
//   q = 0: div    at (i, j, k)          = D+x vx/dx + D+y vy/dy + D-z vz/dz
//   q = 1: curl_x at (i, j-1/2, k+1/2)  = D-y vz/dy - D+z vy/dz
//   q = 2: curl_y at (i-1/2, j, k+1/2)  = D+z vx/dz - D-x vz/dx
//   q = 3: curl_z at (i-1/2, j-1/2, k)  = D-x vy/dx - D-y vx/dy
// The outer R cells get 0.  Read on demand (field access, snapshots), not in
// the time loop, so a plain per-cell stencil through the caches serves.
template <int R>
__global__ void k_decom(const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
                        float* __restrict__ out, int q, int nnx, int nny, int nnz, int px, long pxy, Coef cf,
                        float ix, float iy, float iz) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)nnx * nny * nnz) return;
    const int i = (int)(t % nnx), j = (int)((t / nnx) % nny), k = (int)(t / ((long)nnx * nny));
    const long n = (long)k * pxy + (long)j * px + i;
    if (i < R || i >= nnx - R || j < R || j >= nny - R || k < R || k >= nnz - R) { out[n] = 0.f; return; }
    auto dp = [&](const float* f, long s) {
        float a = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) a += cf.c[m] * (f[n + m * s] - f[n - (m - 1) * s]);
        return a;
    };
    auto dm = [&](const float* f, long s) {
        float a = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) a += cf.c[m] * (f[n + (m - 1) * s] - f[n - m * s]);
        return a;
    };
    float v;
    if (q == 0) v = ix * dp(vx, 1) + iy * dp(vy, px) + iz * dm(vz, pxy);
    else if (q == 1) v = iy * dm(vz, px) - iz * dp(vy, pxy);
    else if (q == 2) v = iz * dp(vx, pxy) - ix * dm(vz, 1);
    else v = ix * dm(vy, 1) - iy * dm(vx, px);
    out[n] = v;
}

// Plain streaming copy (float4), the calibration kernel for HBM counters and
// the practical bandwidth ceiling: reads and writes n4*16 bytes.
[[maybe_unused]] static __global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, long n4) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n4; t += (long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

// Fill kernels (used instead of hipMemsetAsync so the step can be graph-captured).
[[maybe_unused]] static __global__ void k_fill_f32(float* __restrict__ p, long n, float v) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = v;
}
[[maybe_unused]] static __global__ void k_fill_i32(int* __restrict__ p, long n, int v) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = v;
}

}  // namespace erm
