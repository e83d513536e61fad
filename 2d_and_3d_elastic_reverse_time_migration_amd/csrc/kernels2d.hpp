// kernels2d.hpp - gfx950 kernels of the 2D elastic / viscoelastic step, the
// LSRTM Born sources and the exact discrete adjoint.
//
// Layout: z fastest, idx = i*pz + k (the reference's 2D order ix*nz+iz,
// viscoelastic_equation.cu:90-92, with the z pitch rounded up to 32 floats).
// A workgroup is 64 lanes along z (one wavefront per x row, coalesced) by
// TXR rows along x; stencil inputs go through LDS tiles with R-cell halos in
// both directions.  Only interior cells (R <= i < nnx-R, R <= k < nnz-R) are
// updated.  Staggering (viscoelastic_equation.cu:137-176): sxx, szz, tp, r_p,
// r_xx, r_zz at (i,k); vx (i-1/2,k); vz (i,k+1/2); sxz, r_xz (i-1/2,k+1/2).
#pragma once
#include <hip/hip_runtime.h>

namespace erm2 {

constexpr int TZ = 64;  // lanes along z

struct Coef2 {
    float c[7];
};

struct G2 {
    int nnx, nnz, pz;
    float cx, cz, dtr;   // dt/dx, dt/dz (dt in s), dt
    const float* att;    // damping d per cell
    int bl, bu, nx, nz;  // model region
    float qsign;         // sign of the memory-variable term in the stress update (-1: Q-compensated)
    float dsgn;          // sign of the derivative terms (-1: reversed propagation, rfwd_..._old)
    // C-PML (no reference counterpart; null: damping boundary).  psi: 10
    // memory variables of ncell floats each -- stress side 0 Dx vx (x full),
    // 1 Dx vz (x half), 2 Dz vz (z full), 3 Dz vx (z half); velocity side
    // 4 Dx sxx (x half), 5 Dz sxz (z full), 6 Dx sxz (x full), 7 Dz szz
    // (z half), 8 Dx tp (x half), 9 Dz tp (z half).  pcx [4][nnx] and pcz
    // [4][nnz] hold (a_full, b_full, a_half, b_half); half = x - 1/2, z + 1/2.
    float* psi;
    const float* pcx;
    const float* pcz;
    long ncell;
    // Batched shots (LSRTM operators): shot blockIdx.z of a launch works on
    // its own copy of every wavefield array at + z*sst floats, of the E/F
    // history at + z*hst and of the gradient at + z*gst, with its source cell
    // srcs[z]; the model is shared.  Zero strides / null srcs: one shot.
    long sst, hst, gst;
    const long* srcs;
};


// e += (psi = b psi + a e) for memory variable q at cell n, axis coefficients
// pc of length nn at index idx; where a = 0 psi stays 0 and e passes.
__device__ __forceinline__ float pml2(const G2& g, int q, long n, const float* pc, int nn, int half, int idx, float e) {
    const float a = pc[(2 * half) * nn + idx], b = pc[(2 * half + 1) * nn + idx];
    if (a == 0.f) return e;
    float* ps = g.psi + (long)q * g.ncell + n;
    const float v = b * *ps + a * e;
    *ps = v;
    return e + v;
}

struct F2 {  // one wavefield state (forward, Born, or adjoint)
    float *vx, *vz, *xx, *zz, *xz, *tp, *vxp, *vzp;
    float* r[3][4];  // memory variables per mechanism: p, xx, zz, xz
};

// the state arrays of shot z of a batched launch
__device__ __forceinline__ F2 shot_state(F2 f, long off) {
    if (off == 0) return f;
    f.vx += off; f.vz += off; f.xx += off; f.zz += off; f.xz += off; f.tp += off; f.vxp += off; f.vzp += off;
    for (int l = 0; l < 3; ++l)
        for (int q = 0; q < 4; ++q)
            if (f.r[l][q]) f.r[l][q] += off;
    return f;
}

struct M2 {  // model: unrelaxed moduli, buoyancy, SLS coefficients, loaded vp / vs
    const float *M, *U, *B;
    const float *vp, *vs;
    const float* kP[3];
    const float* kS[3];
    const float* decay[3];  // dt / tau_sigma_l
};

struct Ex2 {
    int* time;
    float *amp, *tp, *vxp, *vzp, *img_tp, *img_pp, *img_ps;
    float *rvxp, *rvzp;  // receiver vector P at the imaging step (angle gathers)
    int* tp_time;        // tp excitation: step of max |tp| (minus tp_shift) and the signed tp there
    float* tp_amp;
    int tp_shift;
};

__device__ __forceinline__ int clamp2(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// LDS tile [TXR+2R][TZ+2R] of `f` around the block's (i0, k0) corner; lanes
// load the centre, the z halo (lanes k<R and k>=TZ-R) and the x halo rows.
template <int R, int TXR>
__device__ __forceinline__ void load_tile(float (*t)[TZ + 2 * R], const float* __restrict__ f, const G2& g, int i0,
                                          int k0) {
    const int tz = threadIdx.x, tx = threadIdx.y;
    const int i = clamp2(i0 + tx, 0, g.nnx - 1);
    const int k = clamp2(k0 + tz, 0, g.nnz - 1);
    t[tx + R][tz + R] = f[(long)i * g.pz + k];
    if (tz < R) t[tx + R][tz] = f[(long)i * g.pz + clamp2(k0 - R + tz, 0, g.nnz - 1)];
    if (tz >= TZ - R) t[tx + R][tz + 2 * R] = f[(long)i * g.pz + clamp2(k0 + tz + R, 0, g.nnz - 1)];
    for (int r = tx; r < 2 * R; r += TXR) {
        const int ii = clamp2(r < R ? i0 - R + r : i0 + TXR + (r - R), 0, g.nnx - 1);
        t[r < R ? r : TXR + r][tz + R] = f[(long)ii * g.pz + k];
    }
}

// D+ f centred at +1/2 and D- f at -1/2 along x (rows) or z (columns) of an LDS tile
template <int R>
__device__ __forceinline__ float dpx(const float (*t)[TZ + 2 * R], const float* c, int a, int b) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= R; ++m) s += c[m] * (t[a + m][b] - t[a - m + 1][b]);
    return s;
}
template <int R>
__device__ __forceinline__ float dmx(const float (*t)[TZ + 2 * R], const float* c, int a, int b) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= R; ++m) s += c[m] * (t[a + m - 1][b] - t[a - m][b]);
    return s;
}
template <int R>
__device__ __forceinline__ float dpz(const float (*t)[TZ + 2 * R], const float* c, int a, int b) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= R; ++m) s += c[m] * (t[a][b + m] - t[a][b - m + 1]);
    return s;
}
template <int R>
__device__ __forceinline__ float dmz(const float (*t)[TZ + 2 * R], const float* c, int a, int b) {
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= R; ++m) s += c[m] * (t[a][b + m - 1] - t[a][b - m]);
    return s;
}

__device__ __forceinline__ bool in_model(const G2& g, int i, int k, long* q) {
    if (i < g.bl || i >= g.bl + g.nx || k < g.bu || k >= g.bu + g.nz) return false;
    *q = (long)(i - g.bl) * g.nz + (k - g.bu);
    return true;
}

// ---------------------------------------------------------------------------
// Stress + memory + tp (fwd_txxzzxzpp_viscoelastic_and_memory_3parameterization,
// viscoelastic_equation.cu:184-325; tp as :172), L mechanisms, explosive
// source fused (src = cell index or -1).  With SAVE, the strain increments
// E = (Vx, Vz, S) of the model region go to Eh ([3][nx*nz]) for LSRTM.
// ---------------------------------------------------------------------------
template <int R, int L, int TXR, bool SAVE>
__global__ __launch_bounds__(TZ* TXR) void k2_stress(F2 f, M2 m, G2 g, Coef2 cf, long src, float srcval,
                                                     float* __restrict__ Eh) {
    __shared__ float tvx[TXR + 2 * R][TZ + 2 * R];
    __shared__ float tvz[TXR + 2 * R][TZ + 2 * R];
    if (const int z = blockIdx.z) {
        f = shot_state(f, z * g.sst);
        if constexpr (SAVE) Eh += z * g.hst;
    }
    if (src >= 0 && g.srcs) src = g.srcs[blockIdx.z];
    const int k0 = blockIdx.x * TZ, i0 = blockIdx.y * TXR;
    load_tile<R, TXR>(tvx, f.vx, g, i0, k0);
    load_tile<R, TXR>(tvz, f.vz, g, i0, k0);
    __syncthreads();
    const int k = k0 + threadIdx.x, i = i0 + threadIdx.y;
    if (i < R || i >= g.nnx - R || k < R || k >= g.nnz - R) return;
    float c[R + 1];
#pragma unroll
    for (int q = 1; q <= R; ++q) c[q] = cf.c[q];
    const int a_ = threadIdx.y + R, b_ = threadIdx.x + R;
    const long n = (long)i * g.pz + k;
    float E0, E1, E2;
    if (!g.psi) {
        E0 = g.cx * dpx<R>(tvx, c, a_, b_);
        E1 = g.cz * dmz<R>(tvz, c, a_, b_);
        E2 = g.cx * dmx<R>(tvz, c, a_, b_) + g.cz * dpz<R>(tvx, c, a_, b_);
    } else {
        E0 = pml2(g, 0, n, g.pcx, g.nnx, 0, i, g.cx * dpx<R>(tvx, c, a_, b_));
        E1 = pml2(g, 2, n, g.pcz, g.nnz, 0, k, g.cz * dmz<R>(tvz, c, a_, b_));
        E2 = pml2(g, 1, n, g.pcx, g.nnx, 1, i, g.cx * dmx<R>(tvz, c, a_, b_)) +
             pml2(g, 3, n, g.pcz, g.nnz, 1, k, g.cz * dpz<R>(tvx, c, a_, b_));
    }
    if constexpr (SAVE) {
        long q;
        if (in_model(g, i, k, &q)) {
            const long mn = (long)g.nx * g.nz;
            Eh[q] = E0; Eh[mn + q] = E1; Eh[2 * mn + q] = E2;
        }
    }
    const float h = g.att[n] * g.dtr * 0.5f;
    const float a = 1.0f / (1.0f + h), b = 1.0f - h;
    const float M = m.M[n], U = m.U[n], La = M - 2.0f * U;
    float sp = 0.f, sxx = 0.f, szz = 0.f, sxz = 0.f;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const float bl = b - m.decay[l][n], kP = m.kP[l][n], kS = m.kS[l][n];
        const float rp = a * (bl * f.r[l][0][n] + kP * (E0 + E1));
        const float rxx = a * (bl * f.r[l][1][n] - 2.0f * kS * E1);
        const float rzz = a * (bl * f.r[l][2][n] - 2.0f * kS * E0);
        const float rxz = a * (bl * f.r[l][3][n] + kS * E2);
        f.r[l][0][n] = rp; f.r[l][1][n] = rxx; f.r[l][2][n] = rzz; f.r[l][3][n] = rxz;
        sp += rp; sxx += rxx; szz += rzz; sxz += rxz;
    }
    float xx = f.xx[n], zz = f.zz[n];
    if (n == src) { xx += srcval; zz += srcval; }
    // memory term dt * r; with qsign = -1 its sign is flipped (Q-compensated
    // propagation, rfwd_txxzzxzpp_viscoelastic_and_memory, viscoelastic_equation.cu:1541-1569)
    const float dq = g.dtr * g.qsign;
    // dsgn = -1 (with qsign = -1): the reversed step of
    // rfwd_txxzzxzpp_viscoelastic_and_memory_old (:1407-1433); the memory
    // variables above are driven by +E as in that kernel (an exact sign flip,
    // so the forward step is unchanged bit for bit)
    const float e0 = g.dsgn * E0, e1 = g.dsgn * E1, e2 = g.dsgn * E2;
    f.xx[n] = a * (b * xx + M * e0 + La * e1 + dq * (sp + sxx));
    f.zz[n] = a * (b * zz + La * e0 + M * e1 + dq * (sp + szz));
    f.tp[n] = a * (b * f.tp[n] + M * (e0 + e1) + dq * sp);
    f.xz[n] = a * (b * f.xz[n] + U * e2 + dq * sxz);
}

// ---------------------------------------------------------------------------
// Velocity + vector P (the 2D counterpart of fwd_vx/vz and fwd_vxp_vzp),
// MODE 1: excitation (caculate_ex_amp_time_new, elastic_2D_kernel_3.cu:30-48),
// MODE 2: imaging (imaging_correlation_ex_2D / imaging_inner_product_ex_2D,
// :542-580).  With SAVE, the stress forces F = (fx, fz) of the model region
// go to Fh ([2][nx*nz]).
// ---------------------------------------------------------------------------
template <int R, int TXR, int MODE, bool SAVE>
__global__ __launch_bounds__(TZ* TXR) void k2_velocity(F2 f, M2 m, G2 g, Coef2 cf, Ex2 ex, int it,
                                                       float* __restrict__ Fh) {
    __shared__ float txx[TXR + 2 * R][TZ + 2 * R];
    __shared__ float tzz[TXR + 2 * R][TZ + 2 * R];
    __shared__ float txz[TXR + 2 * R][TZ + 2 * R];
    __shared__ float ttp[TXR + 2 * R][TZ + 2 * R];
    if (const int z = blockIdx.z) {  // batched shots: MODE 0 only
        f = shot_state(f, z * g.sst);
        if constexpr (SAVE) Fh += z * g.hst;
    }
    const int k0 = blockIdx.x * TZ, i0 = blockIdx.y * TXR;
    load_tile<R, TXR>(txx, f.xx, g, i0, k0);
    load_tile<R, TXR>(tzz, f.zz, g, i0, k0);
    load_tile<R, TXR>(txz, f.xz, g, i0, k0);
    load_tile<R, TXR>(ttp, f.tp, g, i0, k0);
    __syncthreads();
    const int k = k0 + threadIdx.x, i = i0 + threadIdx.y;
    if (i < R || i >= g.nnx - R || k < R || k >= g.nnz - R) return;
    float c[R + 1];
#pragma unroll
    for (int q = 1; q <= R; ++q) c[q] = cf.c[q];
    const int a_ = threadIdx.y + R, b_ = threadIdx.x + R;
    const long n = (long)i * g.pz + k;
    float fx, fz, px, pz;
    if (!g.psi) {
        fx = g.cx * dmx<R>(txx, c, a_, b_) + g.cz * dmz<R>(txz, c, a_, b_);
        fz = g.cx * dpx<R>(txz, c, a_, b_) + g.cz * dpz<R>(tzz, c, a_, b_);
        px = g.cx * dmx<R>(ttp, c, a_, b_);
        pz = g.cz * dpz<R>(ttp, c, a_, b_);
    } else {
        fx = pml2(g, 4, n, g.pcx, g.nnx, 1, i, g.cx * dmx<R>(txx, c, a_, b_)) +
             pml2(g, 5, n, g.pcz, g.nnz, 0, k, g.cz * dmz<R>(txz, c, a_, b_));
        fz = pml2(g, 6, n, g.pcx, g.nnx, 0, i, g.cx * dpx<R>(txz, c, a_, b_)) +
             pml2(g, 7, n, g.pcz, g.nnz, 1, k, g.cz * dpz<R>(tzz, c, a_, b_));
        px = pml2(g, 8, n, g.pcx, g.nnx, 1, i, g.cx * dmx<R>(ttp, c, a_, b_));
        pz = pml2(g, 9, n, g.pcz, g.nnz, 1, k, g.cz * dpz<R>(ttp, c, a_, b_));
    }
    if constexpr (SAVE) {
        long q;
        if (in_model(g, i, k, &q)) {
            const long mn = (long)g.nx * g.nz;
            Fh[q] = fx; Fh[mn + q] = fz;
        }
    }
    const float h = g.att[n] * g.dtr * 0.5f;
    const float a = 1.0f / (1.0f + h), b = 1.0f - h, bo = m.B[n];
    const float nvx = a * (b * f.vx[n] + bo * (g.dsgn * fx));
    const float nvz = a * (b * f.vz[n] + bo * (g.dsgn * fz));
    const float nxp = a * (b * f.vxp[n] + bo * (g.dsgn * px));
    const float nzp = a * (b * f.vzp[n] + bo * (g.dsgn * pz));
    f.vx[n] = nvx; f.vz[n] = nvz; f.vxp[n] = nxp; f.vzp[n] = nzp;
    if constexpr (MODE == 1) {
        const float amp = sqrtf(nxp * nxp + nzp * nzp);
        if (ex.amp[n] < amp) {
            ex.time[n] = it; ex.amp[n] = amp; ex.tp[n] = ttp[a_][b_];
            ex.vxp[n] = nxp; ex.vzp[n] = nzp;
        }
        // caculate_ex_tp_time_new (elastic_2D_kernel_3.cu:50-69): its own step, signed amplitude
        const float tpc = ttp[a_][b_];
        if (fabsf(ex.tp_amp[n]) < fabsf(tpc)) { ex.tp_time[n] = it - ex.tp_shift; ex.tp_amp[n] = tpc; }
    } else if constexpr (MODE == 2) {
        if (ex.tp_time[n] == it) {  // imaging_correlation_ex_2D (:542-558)
            const float ta = ex.tp_amp[n];
            ex.img_tp[n] = ta != 0.f ? ttp[a_][b_] / ta : 0.f;
        }
        if (ex.time[n] == it) {
            const float e1 = ex.vxp[n], e2 = ex.vzp[n], den = e1 * e1 + e2 * e2;
            ex.img_pp[n] = den > 0.f ? (e1 * nxp + e2 * nzp) / den : 0.f;
            ex.img_ps[n] = den > 0.f ? (e1 * (nvx - nxp) + e2 * (nvz - nzp)) / den : 0.f;
            ex.rvxp[n] = nxp; ex.rvzp[n] = nzp;
        }
    }
}

// The _new variants caculate_ex_x_z_new / imaging_inner_product_ex_2D_new
// (elastic_2D_kernel_3.cu:91-109,582-605), run after the velocity update of
// step it (the neighbours' values of the step are complete only then): at
// interior cells whose excitation step is it, vzp is the mean over
// (i, k+-1), (i+-1, k) -- MODE 1 in the stored excitation component, MODE 2 in
// the receiver P and S motion of the PP / PS images.
template <int MODE>
__global__ void k2_ex_avg(F2 f, Ex2 ex, G2 g, int it, int R) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (k < R || k >= g.nnz - R || i < R || i >= g.nnx - R) return;
    const long n = (long)i * g.pz + k;
    if (ex.time[n] != it) return;
    auto avg = [&](const float* a) { return (a[n + 1] + a[n - 1] + a[n + g.pz] + a[n - g.pz]) * 0.25f; };
    if constexpr (MODE == 1) {
        ex.vzp[n] = avg(f.vzp);
    } else {
        const float nxp = f.vxp[n], nzp = avg(f.vzp);
        const float nxs = f.vx[n] - nxp, nzs = avg(f.vz) - nzp;
        const float e1 = ex.vxp[n], e2 = ex.vzp[n], den = e1 * e1 + e2 * e2;
        ex.img_pp[n] = den > 0.f ? (e1 * nxp + e2 * nzp) / den : 0.f;
        ex.img_ps[n] = den > 0.f ? (e1 * nxs + e2 * nzs) / den : 0.f;
    }
}

// ---------------------------------------------------------------------------
// Born secondary sources (cuda_cal_dem_parameter_viscoelastic_media_new with
// inversion_para = 2, viscoelastic_equation.cu:2398-2505), added after the
// stress / velocity updates of the Born state (the updates are linear, so a
// source inside the damping bracket equals a*source added afterwards).
// dm = [3][nx*nz] relative perturbations (dvp/vp, dvs/vs, drho/rho).
// ---------------------------------------------------------------------------
// Relative perturbations (sP, sS) of the P and S moduli rho vp^2, rho vs^2
// from the inverted parameters (m1, m2, m3) of inversion_para (the mapping
// of cuda_cal_dem_parameter_viscoelastic_media_new, viscoelastic_equation.cu:
// 2431-2486, for relative perturbations): 1 (lambda, mu, rho):
// sP = (lambda m1 + 2 mu m2) / (lambda + 2 mu), sS = m2; 2 (vp, vs, rho):
// sP = 2 m1 + m3, sS = 2 m2 + m3; 3 (Ip, Is, rho): sP = 2 m1 - m3,
// sS = 2 m2 - m3.
__device__ __forceinline__ void rel_moduli(int para, float vp, float vs, float m1, float m2, float m3, float* sP,
                                           float* sS) {
    if (para == 1) {
        *sP = ((vp * vp - 2.0f * vs * vs) * m1 + 2.0f * vs * vs * m2) / (vp * vp);
        *sS = m2;
    } else if (para == 3) {
        *sP = 2.0f * m1 - m3;
        *sS = 2.0f * m2 - m3;
    } else {
        *sP = 2.0f * m1 + m3;
        *sS = 2.0f * m2 + m3;
    }
}

template <int L>
__global__ void k2_born_stress_src(F2 d, M2 m, G2 g, const float* __restrict__ dm, const float* __restrict__ Eh,
                                   int para) {
    if (const int z = blockIdx.z) { d = shot_state(d, z * g.sst); Eh += z * g.hst; }
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (k >= g.nnz) return;
    long q;
    if (!in_model(g, i, k, &q)) return;
    const long n = (long)i * g.pz + k, mn = (long)g.nx * g.nz;
    const float h = g.att[n] * g.dtr * 0.5f, a = 1.0f / (1.0f + h);
    const float E0 = Eh[q], E1 = Eh[mn + q], E2 = Eh[2 * mn + q];
    float sP, sS;
    rel_moduli(para, m.vp[n], m.vs[n], dm[q], dm[mn + q], dm[2 * mn + q], &sP, &sS);
    const float dM = m.M[n] * sP, dU = m.U[n] * sS, dL = dM - 2.0f * dU;
    float sp = 0.f, sxx = 0.f, szz = 0.f, sxz = 0.f;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const float dkP = m.kP[l][n] * sP, dkS = m.kS[l][n] * sS;
        const float rp = a * dkP * (E0 + E1), rxx = a * (-2.0f * dkS * E1), rzz = a * (-2.0f * dkS * E0),
                    rxz = a * dkS * E2;
        d.r[l][0][n] += rp; d.r[l][1][n] += rxx; d.r[l][2][n] += rzz; d.r[l][3][n] += rxz;
        sp += rp; sxx += rxx; szz += rzz; sxz += rxz;
    }
    d.xx[n] += a * (dM * E0 + dL * E1 + g.dtr * (sp + sxx));
    d.zz[n] += a * (dL * E0 + dM * E1 + g.dtr * (sp + szz));
    d.xz[n] += a * (dU * E2 + g.dtr * sxz);
}

__global__ void k2_born_velocity_src(F2 d, M2 m, G2 g, const float* __restrict__ dm, const float* __restrict__ Fh) {
    if (const int z = blockIdx.z) { d = shot_state(d, z * g.sst); Fh += z * g.hst; }
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (k >= g.nnz) return;
    long q;
    if (!in_model(g, i, k, &q)) return;
    const long n = (long)i * g.pz + k, mn = (long)g.nx * g.nz;
    const float h = g.att[n] * g.dtr * 0.5f, a = 1.0f / (1.0f + h);
    const float s = -a * m.B[n] * dm[2 * mn + q];  // delta(1/rho) = -(drho/rho)/rho
    d.vx[n] += s * Fh[q];
    d.vz[n] += s * Fh[mn + q];
}

// ---------------------------------------------------------------------------
// Adjoint (exact transpose of one Born step, run backwards in time).
// Transpose of the velocity update: ls += F^T(Z), Z = a*B*lv on interior
// cells (0 elsewhere), with (D-)^T = -D+ and (D+)^T = -D-; the scaling
// lv <- a*b*lv is deferred into k2a_s2.  Density gradient from the stored
// forces F0 of the step.
// ---------------------------------------------------------------------------
template <int R, int TXR>
__global__ __launch_bounds__(TZ* TXR) void k2a_v(F2 l, M2 m, G2 g, Coef2 cf, const float* __restrict__ Fh,
                                                 float* __restrict__ grad) {
    __shared__ float zx[TXR + 2 * R][TZ + 2 * R];
    __shared__ float zz[TXR + 2 * R][TZ + 2 * R];
    if (const int z = blockIdx.z) { l = shot_state(l, z * g.sst); Fh += z * g.hst; grad += z * g.gst; }
    const int k0 = blockIdx.x * TZ, i0 = blockIdx.y * TXR;
    const int tz = threadIdx.x, tx = threadIdx.y;
    auto Zat = [&](int ii, int kk, float* ox, float* oz) {
        if (ii < R || ii >= g.nnx - R || kk < R || kk >= g.nnz - R) { *ox = 0.f; *oz = 0.f; return; }
        const long n = (long)ii * g.pz + kk;
        const float h = g.att[n] * g.dtr * 0.5f, s = m.B[n] / (1.0f + h);
        *ox = s * l.vx[n];
        *oz = s * l.vz[n];
    };
    {
        float a, b;
        Zat(i0 + tx, k0 + tz, &a, &b);
        zx[tx + R][tz + R] = a; zz[tx + R][tz + R] = b;
        if (tz < R) { Zat(i0 + tx, k0 - R + tz, &a, &b); zx[tx + R][tz] = a; zz[tx + R][tz] = b; }
        if (tz >= TZ - R) { Zat(i0 + tx, k0 + tz + R, &a, &b); zx[tx + R][tz + 2 * R] = a; zz[tx + R][tz + 2 * R] = b; }
        for (int r = tx; r < 2 * R; r += TXR) {
            const int ii = r < R ? i0 - R + r : i0 + TXR + (r - R);
            Zat(ii, k0 + tz, &a, &b);
            zx[r < R ? r : TXR + r][tz + R] = a; zz[r < R ? r : TXR + r][tz + R] = b;
        }
    }
    __syncthreads();
    const int k = k0 + tz, i = i0 + tx;
    if (i < R || i >= g.nnx - R || k < R || k >= g.nnz - R) return;
    float c[R + 1];
#pragma unroll
    for (int q = 1; q <= R; ++q) c[q] = cf.c[q];
    const int a_ = tx + R, b_ = tz + R;
    const long n = (long)i * g.pz + k;
    l.xx[n] += -g.cx * dpx<R>(zx, c, a_, b_);
    l.xz[n] += -g.cz * dpz<R>(zx, c, a_, b_) - g.cx * dmx<R>(zz, c, a_, b_);
    l.zz[n] += -g.cz * dmz<R>(zz, c, a_, b_);
    long q;
    if (in_model(g, i, k, &q)) {
        const long mn = (long)g.nx * g.nz;
        grad[2 * mn + q] += -(Fh[q] * zx[a_][b_] + Fh[mn + q] * zz[a_][b_]);
    }
}

// Transpose of the stress/memory update, pointwise part: mu_s = a ls,
// mu_r = lr + dt P^T mu_s, W = C mu_s + a K^T mu_r (to scratch, interior
// cells), ls <- b mu_s, lr <- a (b - dt/tau) mu_r, gradients (mp, ms, mr)
// from the stored strain increments E0 of the step.
template <int L>
__global__ void k2a_s1(F2 l, M2 m, G2 g, const float* __restrict__ Eh, float* __restrict__ grad,
                       float* __restrict__ W0, float* __restrict__ W1, float* __restrict__ W2, int R) {
    if (const int z = blockIdx.z) {
        l = shot_state(l, z * g.sst);
        Eh += z * g.hst; grad += z * g.gst;
        W0 += z * g.sst; W1 += z * g.sst; W2 += z * g.sst;
    }
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (k < R || k >= g.nnz - R || i < R || i >= g.nnx - R) return;
    const long n = (long)i * g.pz + k;
    const float h = g.att[n] * g.dtr * 0.5f, a = 1.0f / (1.0f + h), b = 1.0f - h;
    const float mxx = a * l.xx[n], mzz = a * l.zz[n], mxz = a * l.xz[n];
    const float M = m.M[n], U = m.U[n], La = M - 2.0f * U;
    float wx = M * mxx + La * mzz, wz = La * mxx + M * mzz, ws = U * mxz;
    long q;
    const bool inm = in_model(g, i, k, &q);
    const long mn = (long)g.nx * g.nz;
    float E0 = 0.f, E1 = 0.f, E2 = 0.f;
    if (inm) { E0 = Eh[q]; E1 = Eh[mn + q]; E2 = Eh[2 * mn + q]; }
    float gkp = 0.f, gks = 0.f;
#pragma unroll
    for (int s = 0; s < L; ++s) {
        const float rp = l.r[s][0][n] + g.dtr * (mxx + mzz), rxx = l.r[s][1][n] + g.dtr * mxx;
        const float rzz = l.r[s][2][n] + g.dtr * mzz, rxz = l.r[s][3][n] + g.dtr * mxz;
        const float kP = m.kP[s][n], kS = m.kS[s][n], bl = b - m.decay[s][n];
        wx += a * (kP * rp - 2.0f * kS * rzz);
        wz += a * (kP * rp - 2.0f * kS * rxx);
        ws += a * kS * rxz;
        if (inm) {
            gkp += kP * (E0 + E1) * a * rp;
            gks += kS * (-2.0f * E1 * a * rxx - 2.0f * E0 * a * rzz + E2 * a * rxz);
        }
        l.r[s][0][n] = a * bl * rp; l.r[s][1][n] = a * bl * rxx;
        l.r[s][2][n] = a * bl * rzz; l.r[s][3][n] = a * bl * rxz;
    }
    W0[n] = wx; W1[n] = wz; W2[n] = ws;
    if (inm) {
        const float gM = (E0 + E1) * (mxx + mzz), gU = -2.0f * (E1 * mxx + E0 * mzz) + E2 * mxz;
        grad[q] += 2.0f * M * gM + 2.0f * gkp;
        grad[mn + q] += 2.0f * U * gU + 2.0f * gks;
        grad[2 * mn + q] += M * gM + U * gU + gkp + gks;
    }
    l.xx[n] = b * mxx; l.zz[n] = b * mzz; l.xz[n] = b * mxz;
}

// Transpose of the strain operator: lv <- a b lv (deferred from k2a_v) + E^T W.
template <int R, int TXR>
__global__ __launch_bounds__(TZ* TXR) void k2a_s2(F2 l, G2 g, Coef2 cf, const float* __restrict__ W0,
                                                  const float* __restrict__ W1, const float* __restrict__ W2) {
    __shared__ float w0[TXR + 2 * R][TZ + 2 * R];
    __shared__ float w1[TXR + 2 * R][TZ + 2 * R];
    __shared__ float w2[TXR + 2 * R][TZ + 2 * R];
    if (const int z = blockIdx.z) { l = shot_state(l, z * g.sst); W0 += z * g.sst; W1 += z * g.sst; W2 += z * g.sst; }
    const int k0 = blockIdx.x * TZ, i0 = blockIdx.y * TXR;
    load_tile<R, TXR>(w0, W0, g, i0, k0);
    load_tile<R, TXR>(w1, W1, g, i0, k0);
    load_tile<R, TXR>(w2, W2, g, i0, k0);
    __syncthreads();
    const int k = k0 + threadIdx.x, i = i0 + threadIdx.y;
    if (i < R || i >= g.nnx - R || k < R || k >= g.nnz - R) return;
    float c[R + 1];
#pragma unroll
    for (int q = 1; q <= R; ++q) c[q] = cf.c[q];
    const int a_ = threadIdx.y + R, b_ = threadIdx.x + R;
    const long n = (long)i * g.pz + k;
    const float h = g.att[n] * g.dtr * 0.5f, a = 1.0f / (1.0f + h), b = 1.0f - h;
    l.vx[n] = a * b * l.vx[n] - g.cx * dmx<R>(w0, c, a_, b_) - g.cz * dmz<R>(w2, c, a_, b_);
    l.vz[n] = a * b * l.vz[n] - g.cz * dpz<R>(w1, c, a_, b_) - g.cx * dpx<R>(w2, c, a_, b_);
}

// ---------------------------------------------------------------------------
// Reference-mode adjoint (adjoint_mode 1): the reference's continuous-adjoint
// state equations, stepping backward in time with the forward kernels'
// damping bracket a[b u + ...].  Adjoint state in an F2: vx, vz the adjoint
// velocities, xx, zz, xz the strain-like adjoint stresses, r[l][1..3] the
// memory variables (xx, zz, xz) of mechanism l (r[l][0] unused: the adjoint
// r_p of the reference feeds neither the velocities nor the gradient).
//   k2r_s: stresses += E(v) (the strain increments, no moduli: the kernel
//          that updates the reference's adjoint stresses is not shipped; the
//          positions of viscoelastic_equation.cu:137-176), then per mechanism
//          r = a[(b - dt/tau) r + dt s] (adjoint_fwd_memory_new, :2190-2196).
//   k2r_v: velocities (adjoint_fwd_vx/vz_viscoelastic, :1923-1971,
//          :2071-2120) with the cell's moduli outside the derivatives,
//          Mp = M/rho, Ms = U/rho, P_l = kP_l/rho, S_l = kS_l/rho:
//            vx += Mp cx D-x sxx + (Mp-2Ms) cx D-x szz + Ms cz D-z sxz
//                  + sum_l P_l cx D-x rxx + (P_l-2S_l) cx D-x rzz + S_l cz D-z rxz
//            vz += Mp cz D+z szz + (Mp-2Ms) cz D+z sxx + Ms cx D+x sxz
//                  + sum_l (P_l-2S_l) cz D+z rxx + P_l cz D+z rzz + S_l cx D+x rxz
//          then, at model cells, the gradient of
//          cal_gradient_in_viscoelastic_media_new (:2248-2295) from the
//          background's strain increments E and forces F of the step:
//            g_rho += vx vx_t + vz vz_t,  v_t = B F / dt
//            g_lam -= [mp (sxx+szz) + sum_l np_l (rxx+rzz)] (vx_x + vz_z)
//            g_mu  -= 2A + 2B + C + sum_l (2D_l + 2E_l + F_l)
//          with vx_x = E0/dt, vz_z = E1/dt, vx_z + vz_x = E2/dt,
//          A = (mp vx_x + (mp-ms) vz_z) sxx, B = (mp vz_z + (mp-ms) vx_x) szz,
//          C = ms (vx_z+vz_x) sxz, D/E/F likewise with np_l, ns_l and r_l;
//          mp = M/(rho vp^2), ms = U/(rho vs^2), np_l = kP_l/(rho vp^2),
//          ns_l = kS_l/(rho vs^2) (ms = ns_l = 0 where vs = 0).
// grad = [3][nx*nz] (g_lam, g_mu, g_rho), accumulated over the steps.
// ---------------------------------------------------------------------------
template <int R, int L, int TXR>
__global__ __launch_bounds__(TZ* TXR) void k2r_s(F2 l, M2 m, G2 g, Coef2 cf) {
    __shared__ float tvx[TXR + 2 * R][TZ + 2 * R];
    __shared__ float tvz[TXR + 2 * R][TZ + 2 * R];
    if (const int z = blockIdx.z) l = shot_state(l, z * g.sst);
    const int k0 = blockIdx.x * TZ, i0 = blockIdx.y * TXR;
    load_tile<R, TXR>(tvx, l.vx, g, i0, k0);
    load_tile<R, TXR>(tvz, l.vz, g, i0, k0);
    __syncthreads();
    const int k = k0 + threadIdx.x, i = i0 + threadIdx.y;
    if (i < R || i >= g.nnx - R || k < R || k >= g.nnz - R) return;
    float c[R + 1];
#pragma unroll
    for (int q = 1; q <= R; ++q) c[q] = cf.c[q];
    const int a_ = threadIdx.y + R, b_ = threadIdx.x + R;
    const long n = (long)i * g.pz + k;
    const float E0 = g.cx * dpx<R>(tvx, c, a_, b_);
    const float E1 = g.cz * dmz<R>(tvz, c, a_, b_);
    const float E2 = g.cx * dmx<R>(tvz, c, a_, b_) + g.cz * dpz<R>(tvx, c, a_, b_);
    const float h = g.att[n] * g.dtr * 0.5f, a = 1.0f / (1.0f + h), b = 1.0f - h;
    const float sxx = a * (b * l.xx[n] + E0), szz = a * (b * l.zz[n] + E1), sxz = a * (b * l.xz[n] + E2);
    l.xx[n] = sxx; l.zz[n] = szz; l.xz[n] = sxz;
#pragma unroll
    for (int s = 0; s < L; ++s) {
        const float bl = b - m.decay[s][n];
        l.r[s][1][n] = a * (bl * l.r[s][1][n] + g.dtr * sxx);
        l.r[s][2][n] = a * (bl * l.r[s][2][n] + g.dtr * szz);
        l.r[s][3][n] = a * (bl * l.r[s][3][n] + g.dtr * sxz);
    }
}

template <int R, int L, int TXR>
__global__ __launch_bounds__(TZ* TXR) void k2r_v(F2 l, M2 m, G2 g, Coef2 cf, const float* __restrict__ Eh,
                                                 float* __restrict__ grad) {
    __shared__ float t0[TXR + 2 * R][TZ + 2 * R];
    __shared__ float t1[TXR + 2 * R][TZ + 2 * R];
    __shared__ float t2[TXR + 2 * R][TZ + 2 * R];
    if (const int z = blockIdx.z) { l = shot_state(l, z * g.sst); Eh += z * g.hst; grad += z * g.gst; }
    const int k0 = blockIdx.x * TZ, i0 = blockIdx.y * TXR;
    const int k = k0 + threadIdx.x, i = i0 + threadIdx.y;
    const bool upd = !(i < R || i >= g.nnx - R || k < R || k >= g.nnz - R);
    const long n = (long)clamp2(i, 0, g.nnx - 1) * g.pz + clamp2(k, 0, g.nnz - 1);
    float c[R + 1];
#pragma unroll
    for (int q = 1; q <= R; ++q) c[q] = cf.c[q];
    const int a_ = threadIdx.y + R, b_ = threadIdx.x + R;
    const float Bn = m.B[n], Mp = m.M[n] * Bn, Ms = m.U[n] * Bn;
    load_tile<R, TXR>(t0, l.xx, g, i0, k0);
    load_tile<R, TXR>(t1, l.zz, g, i0, k0);
    load_tile<R, TXR>(t2, l.xz, g, i0, k0);
    __syncthreads();
    float fx = Mp * g.cx * dmx<R>(t0, c, a_, b_) + (Mp - 2.0f * Ms) * g.cx * dmx<R>(t1, c, a_, b_) +
               Ms * g.cz * dmz<R>(t2, c, a_, b_);
    float fz = Mp * g.cz * dpz<R>(t1, c, a_, b_) + (Mp - 2.0f * Ms) * g.cz * dpz<R>(t0, c, a_, b_) +
               Ms * g.cx * dpx<R>(t2, c, a_, b_);
    const float sxx = t0[a_][b_], szz = t1[a_][b_], sxz = t2[a_][b_];
    float rsum_l = 0.f, rdef = 0.f;  // memory parts of the gradient (accumulated per mechanism)
    long q = -1;
    const bool inm = upd && in_model(g, i, k, &q);
    const long mn = (long)g.nx * g.nz;
    float vxx = 0.f, vzz = 0.f, vxz = 0.f;
    const float idt = 1.0f / g.dtr;
    if (inm) { vxx = Eh[q] * idt; vzz = Eh[mn + q] * idt; vxz = Eh[2 * mn + q] * idt; }
    const float rho = 1.0f / Bn, vp = m.vp[n], vs = m.vs[n];
    const float ip = 1.0f / (rho * vp * vp), is = vs > 0.f ? 1.0f / (rho * vs * vs) : 0.f;
#pragma unroll
    for (int s = 0; s < L; ++s) {
        __syncthreads();  // the previous tiles are consumed
        load_tile<R, TXR>(t0, l.r[s][1], g, i0, k0);
        load_tile<R, TXR>(t1, l.r[s][2], g, i0, k0);
        load_tile<R, TXR>(t2, l.r[s][3], g, i0, k0);
        __syncthreads();
        const float P = m.kP[s][n] * Bn, S = m.kS[s][n] * Bn;
        fx += P * g.cx * dmx<R>(t0, c, a_, b_) + (P - 2.0f * S) * g.cx * dmx<R>(t1, c, a_, b_) +
              S * g.cz * dmz<R>(t2, c, a_, b_);
        fz += (P - 2.0f * S) * g.cz * dpz<R>(t0, c, a_, b_) + P * g.cz * dpz<R>(t1, c, a_, b_) +
              S * g.cx * dpx<R>(t2, c, a_, b_);
        const float np = m.kP[s][n] * ip, ns = m.kS[s][n] * is;
        const float rxx = t0[a_][b_], rzz = t1[a_][b_], rxz = t2[a_][b_];
        rsum_l += np * (rxx + rzz);
        rdef += 2.0f * (np * vxx + (np - ns) * vzz) * rxx + 2.0f * (np * vzz + (np - ns) * vxx) * rzz + ns * vxz * rxz;
    }
    if (!upd) return;
    const float h = g.att[n] * g.dtr * 0.5f, a = 1.0f / (1.0f + h), b = 1.0f - h;
    const float nvx = a * (b * l.vx[n] + fx), nvz = a * (b * l.vz[n] + fz);
    l.vx[n] = nvx; l.vz[n] = nvz;
    if (inm) {
        const float* Fh = Eh + 3 * mn;
        const float vxt = Bn * Fh[q] * idt, vzt = Bn * Fh[mn + q] * idt;
        const float mp = m.M[n] * ip, ms = m.U[n] * is;
        const float A = (mp * vxx + (mp - ms) * vzz) * sxx, B = (mp * vzz + (mp - ms) * vxx) * szz;
        const float C = ms * vxz * sxz;
        grad[q] -= (mp * (sxx + szz) + rsum_l) * (vxx + vzz);
        grad[mn + q] -= 2.0f * A + 2.0f * B + C + rdef;
        grad[2 * mn + q] += nvx * vxt + nvz * vzt;
    }
}

// Reference-mode gradient (g_lam, g_mu, g_rho) for absolute perturbations,
// mapped to the inverted parameters (relative perturbations) by the
// transpose of the reference's parameter map (cuda_cal_dem_parameter_
// viscoelastic_media_new, viscoelastic_equation.cu:2431-2486), with the sign
// of the density term reversed (with the reference's sign the density part
// of <dm, g> is -<L dm, d>; the moduli parts are +):
//   1: (lam g_lam, mu g_mu, -rho g_rho)
//   2: (2 rho vp^2 g_lam, rho vs^2 (2 g_mu - 4 g_lam), rho (vp^2-2vs^2) g_lam + rho vs^2 g_mu - rho g_rho)
//   3: as 2 with the third component -rho (vp^2-2vs^2) g_lam - rho vs^2 g_mu - rho g_rho.
__global__ void k2r_grad_map(float* __restrict__ grad, M2 m, G2 g, int para) {
    const long mn = (long)g.nx * g.nz;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < mn; q += (long)gridDim.x * blockDim.x) {
        const long n = (long)((int)(q / g.nz) + g.bl) * g.pz + (int)(q % g.nz) + g.bu;
        const float rho = 1.0f / m.B[n], vp = m.vp[n], vs = m.vs[n];
        const float gl = grad[q], gm = grad[mn + q], gr = grad[2 * mn + q];
        const float lam = rho * (vp * vp - 2.0f * vs * vs), mu = rho * vs * vs;
        if (para == 1) {
            grad[q] = lam * gl;
            grad[mn + q] = mu * gm;
            grad[2 * mn + q] = -rho * gr;
        } else {
            grad[q] = 2.0f * rho * vp * vp * gl;
            grad[mn + q] = mu * (2.0f * gm - 4.0f * gl);
            grad[2 * mn + q] = (para == 3 ? -1.0f : 1.0f) * (lam * gl + mu * gm) - rho * gr;
        }
    }
}

// ---------------------------------------------------------------------------
// Cross-correlation RTM (erm2d_rtm_xcorr).  Per receiver step, over interior
// cells, with the source fields (sxp, szp, sdv: vector P and div v after the
// matching source step) taken over the model region only (history layout
// [mn] when hist, else the padded grid) and the receiver's v, vP, div, curl:
//   img [5][mn]: PP += vP_s.vP_r, PS += vP_s.(v_r - vP_r), PPd += div_s div_r,
//                PSc += div_s curl_r (model cells); PScp in k2x_polarity;
//   vote: 1 / 2 / 3 for set > 0 / < 0 / = 0, set = -sgn(vzp_s) sgn(rvxs)
//         (vP_s . rvs), the cases of set_sign_basedon_polarization_ps
//         (elastic_2D_kernel_2.cu:941-982; the case vzp_s < 0, rvxs = 0
//         left undefined there votes 3); prod = div_s curl_r.
// ---------------------------------------------------------------------------
__global__ void k2x_image(G2 g, const float* __restrict__ sxp, const float* __restrict__ szp,
                          const float* __restrict__ sdv, int hist, F2 r, const float* __restrict__ rdv,
                          const float* __restrict__ rcu, float* __restrict__ img, int* __restrict__ vote,
                          float* __restrict__ prod, int R) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (k < R || k >= g.nnz - R || i < R || i >= g.nnx - R) return;
    const long n = (long)i * g.pz + k;
    long q = 0;
    const bool inm = in_model(g, i, k, &q);
    float a = 0.f, b = 0.f, d = 0.f;
    if (inm) {
        const long s = hist ? q : n;
        a = sxp[s]; b = szp[s]; d = sdv[s];
    }
    const float rxs = r.vx[n] - r.vxp[n], rzs = r.vz[n] - r.vzp[n];
    const float dot = a * rxs + b * rzs;
    float set = 0.f;
    if (b >= 0.f && rxs >= 0.f) set = -dot;
    else if (b >= 0.f && rxs < 0.f) set = dot;
    else if (b < 0.f && rxs > 0.f) set = dot;
    else if (b < 0.f && rxs < 0.f) set = -dot;
    vote[n] = set > 0.f ? 1 : (set < 0.f ? 2 : 3);
    const float pr = d * rcu[n];
    prod[n] = pr;
    if (inm) {
        const long mn = (long)g.nx * g.nz;
        img[q] += a * r.vxp[n] + b * r.vzp[n];
        img[mn + q] += dot;
        img[2 * mn + q] += d * rdv[n];
        img[3 * mn + q] += pr;
    }
}
// PS polarity: the votes of the 9x9 window (filter_sign_new_share,
// elastic_2D_kernel_2.cu:875-919; interior cells only) decide each cell's sign
// (compare_sign, :921-939: +1 if its count is not below either other, -1 or 0
// if strictly above both others, otherwise the previous sign stays), then
// PScp += sign * prod at model cells.  One workgroup: 64 cells along z x 4
// rows, the votes of the tile and a 4-cell halo staged in LDS.
constexpr int PW = 4;  // window half width
__global__ __launch_bounds__(TZ * 4) void k2x_polarity(G2 g, const int* __restrict__ vote, int* __restrict__ sign,
                                                       const float* __restrict__ prod, float* __restrict__ img, int R) {
    __shared__ int tv[4 + 2 * PW][TZ + 2 * PW];
    const int tz = threadIdx.x, tx = threadIdx.y;
    const int k0 = blockIdx.x * TZ, i0 = blockIdx.y * 4;
    for (int rr = tx; rr < 4 + 2 * PW; rr += 4)
        for (int cc = tz; cc < TZ + 2 * PW; cc += TZ) {
            const int ii = i0 - PW + rr, kk = k0 - PW + cc;
            const bool in = ii >= R && ii < g.nnx - R && kk >= R && kk < g.nnz - R;
            tv[rr][cc] = in ? vote[(long)ii * g.pz + kk] : 0;
        }
    __syncthreads();
    const int k = k0 + tz, i = i0 + tx;
    if (k < R || k >= g.nnz - R || i < R || i >= g.nnx - R) return;
    int c1 = 0, c2 = 0, c3 = 0;
    for (int di = 0; di <= 2 * PW; ++di)
#pragma unroll
        for (int dk = 0; dk <= 2 * PW; ++dk) {
            const int v = tv[tx + di][tz + dk];
            c1 += v == 1; c2 += v == 2; c3 += v == 3;
        }
    const long n = (long)i * g.pz + k;
    int s = sign[n];
    if (c1 >= c2 && c1 >= c3) s = 1;
    else if (c2 > c1 && c2 > c3) s = -1;
    else if (c3 > c1 && c3 > c2) s = 0;
    sign[n] = s;
    long q;
    if (in_model(g, i, k, &q)) img[4 * (long)g.nx * g.nz + q] += (float)s * prod[n];
}
// source history of one step: vxp, vzp and div of the model region -> h [3][mn]
__global__ void k2x_save_hist(G2 g, const float* __restrict__ vxp, const float* __restrict__ vzp,
                              const float* __restrict__ dv, float* __restrict__ h) {
    const long mn = (long)g.nx * g.nz;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < mn; q += (long)gridDim.x * blockDim.x) {
        const long n = (long)((int)(q / g.nz) + g.bl) * g.pz + (int)(q % g.nz) + g.bu;
        h[q] = vxp[n]; h[mn + q] = vzp[n]; h[2 * mn + q] = dv[n];
    }
}
// boundary strips: the four fields f0..f3 at the ring cells idx[0..nr) to
// buf [4][nr] (save) or back (restore); restore with a null buf writes zeros.
// Batched shots along blockIdx.y: shot z's fields at + z*sst, its buf at + z*bst.
__global__ void k2x_ring(float* __restrict__ f0, float* __restrict__ f1, float* __restrict__ f2, float* __restrict__ f3,
                         const long* __restrict__ idx, long nr, float* __restrict__ buf, int save, long sst, long bst) {
    if (const long z = blockIdx.y) {
        f0 += z * sst; f1 += z * sst; f2 += z * sst; f3 += z * sst;
        if (buf) buf += z * bst;
    }
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nr; t += (long)gridDim.x * blockDim.x) {
        const long n = idx[t];
        if (save) {
            buf[t] = f0[n]; buf[nr + t] = f1[n]; buf[2 * nr + t] = f2[n]; buf[3 * nr + t] = f3[n];
        } else if (buf) {
            f0[n] = buf[t]; f1[n] = buf[nr + t]; f2[n] = buf[2 * nr + t]; f3[n] = buf[3 * nr + t];
        } else {
            f0[n] = 0.f; f1[n] = 0.f; f2[n] = 0.f; f3[n] = 0.f;
        }
    }
}

// Residual records of the mapped reference-mode adjoint: each receiver's
// trace times dt / rho at the receiver cell (the scale that makes the mapped
// gradient an approximate transpose of the Born operator); rec [2][lt][nr]
// per shot, shots along blockIdx.y (record stride rst).
__global__ void k2r_scale_rec(float* __restrict__ rec, const long* __restrict__ idx, int nr, long lt,
                              const float* __restrict__ B, float dt, long rst) {
    rec += blockIdx.y * rst;
    const long n = 2 * lt * nr;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        rec[t] *= dt * B[idx[t % nr]];
}

// Gradient in the inverted parameters: the adjoint accumulates (2 GP, 2 GS,
// GP + GS + GD) in the (vp, vs, rho) parameters, GP / GS the gradients of the
// relative P / S modulus perturbations and GD the direct density part; the
// chain rule of rel_moduli gives parameterisations 1 and 3.
__global__ void k2_grad_para(float* __restrict__ grad, M2 m, G2 g, int para) {
    const long mn = (long)g.nx * g.nz;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < mn; q += (long)gridDim.x * blockDim.x) {
        const float GP = 0.5f * grad[q], GS = 0.5f * grad[mn + q], GD = grad[2 * mn + q] - GP - GS;
        if (para == 1) {
            const long n = (long)((int)(q / g.nz) + g.bl) * g.pz + (int)(q % g.nz) + g.bu;
            const float a = m.vp[n], b = m.vs[n];
            grad[q] = GP * (a * a - 2.0f * b * b) / (a * a);
            grad[mn + q] = GP * 2.0f * b * b / (a * a) + GS;
            grad[2 * mn + q] = GD;
        } else if (para == 3) {
            grad[2 * mn + q] = GD - GP - GS;
        }
    }
}

// out = sum over the ns shots of a batched launch of their gradients (stride n)
__global__ void k2_sum_shots(const float* __restrict__ g, int ns, long n, float* __restrict__ out) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        float a = 0.f;
        for (int z = 0; z < ns; ++z) a += g[z * n + q];
        out[q] = a;
    }
}

// div / curl decomposition (decom, elastic_2D_kernel_2.cu:621-666); with
// svp / svs the decom_new scaling (:671-791): div times vp at the stress node,
// curl times the mean of vs over (i,k), (i,k+1), (i+1,k), (i+1,k+1).
template <int R, int TXR>
__global__ __launch_bounds__(TZ* TXR) void k2_decom(const float* __restrict__ vx, const float* __restrict__ vz,
                                                    float* __restrict__ dv, float* __restrict__ cu, G2 g, Coef2 cf,
                                                    float idx, float idz, const float* __restrict__ svp,
                                                    const float* __restrict__ svs) {
    __shared__ float tvx[TXR + 2 * R][TZ + 2 * R];
    __shared__ float tvz[TXR + 2 * R][TZ + 2 * R];
    const int k0 = blockIdx.x * TZ, i0 = blockIdx.y * TXR;
    load_tile<R, TXR>(tvx, vx, g, i0, k0);
    load_tile<R, TXR>(tvz, vz, g, i0, k0);
    __syncthreads();
    const int k = k0 + threadIdx.x, i = i0 + threadIdx.y;
    if (i >= g.nnx || k >= g.nnz) return;
    const long n = (long)i * g.pz + k;
    if (i < R || i >= g.nnx - R || k < R || k >= g.nnz - R) { dv[n] = 0.f; cu[n] = 0.f; return; }
    float c[R + 1];
#pragma unroll
    for (int q = 1; q <= R; ++q) c[q] = cf.c[q];
    const int a_ = threadIdx.y + R, b_ = threadIdx.x + R;
    float d = idx * dpx<R>(tvx, c, a_, b_) + idz * dmz<R>(tvz, c, a_, b_);
    float r = idz * dpz<R>(tvx, c, a_, b_) - idx * dmx<R>(tvz, c, a_, b_);
    if (svp) d *= svp[n];
    if (svs) r *= (svs[n] + svs[n + 1] + svs[n + g.pz] + svs[n + 1 + g.pz]) * 0.25f;
    dv[n] = d;
    cu[n] = r;
}

// receivers: rec [2][lt][nrec]; batched shots along blockIdx.y (state
// stride sst, record stride rst)
__global__ void k2_record(const float* __restrict__ vx, const float* __restrict__ vz, const long* __restrict__ idx,
                          int nr, float* __restrict__ rec, long comp, long off, long sst, long rst) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    vx += blockIdx.y * sst; vz += blockIdx.y * sst; rec += blockIdx.y * rst;
    rec[off + r] = vx[idx[r]];
    rec[comp + off + r] = vz[idx[r]];
}
__global__ void k2_inject(float* __restrict__ vx, float* __restrict__ vz, const long* __restrict__ idx, int nr,
                          const float* __restrict__ rec, long comp, long off, long sst, long rst) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    vx += blockIdx.y * sst; vz += blockIdx.y * sst; rec += blockIdx.y * rst;
    atomicAdd(&vx[idx[r]], rec[off + r]);
    atomicAdd(&vz[idx[r]], rec[comp + off + r]);
}
// source illumination: W += vx^2 + vz^2 at every grid point (each component
// at its own staggered position)
__global__ void k2_illum_acc(const float* __restrict__ vx, const float* __restrict__ vz, float* __restrict__ W,
                             long n) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const float a = vx[t], b = vz[t];
        W[t] += a * a + b * b;
    }
}
// Angle-domain imaging of the excitation images.  The reference has the
// pieces -- P Poynting directions, least-squares angles over a 5x5 window
// (caculate_ex_angle_pp_only_RTM, elastic_2D_kernel_3.cu:132-208), opening
// angles (caculate_angle_open, elastic_2D_kernel_2.cu:1236-1256), angle bins
// (set_adcigs_pp_new / _ps_new, :1751-1819) -- but no driver calls them; here,
// per model cell with an excitation time:
//   source incidence  ts = atan(sum Sx Sz / sum Sz^2), S = -tp vP at excitation
//                     (Poynting vector of the P wave), 5x5 window, degrees from
//                     the vertical (z down), +ve toward +x;
//   receiver          tr = atan(sum rvxp rvzp / sum rvzp^2) over the same
//                     window, the receiver P motion at each cell's imaging step
//                     (as caculate_ex_angle_rp_only_RTM, :209-285; sign-free:
//                     P motion is along the ray);
//   reflection angle  g = (ts - tr) / 2 (the bisector of the two rays is the
//                     reflector normal);
//   gathers           bin floor((g + nang dang / 2) / dang) of pp / ps
//                     [nang][nx][nz] receives img_pp / img_ps;
//   compensation      the reference's angle-dependent compensation with the
//                     opening angle 2g (cuda_ex_com_pp_ps_sign and
//                     imaging_pp/ps_compensate_dependent_angle_2D_new,
//                     elastic_2D_kernel_3.cu:698-770):
//                     ppc = s(cos 2g) img_pp / (|cos 2g| + 0.03), s(x) = +1 for
//                     x > 0 else -1; psc = img_ps / (|sin 2g| + 0.1) (the
//                     reference also negates both; the images keep their sign).
//                     The PS image of this vector inner-product condition does
//                     not reverse polarity across the source (both factors are
//                     true vectors), so no sign correction is applied.
// Cells without an excitation time get g = 0, no gather entry and 0.
__device__ __forceinline__ float deg_atan_ratio(float num, float den) {
    if (den != 0.f) return atanf(num / den) * (180.f / 3.14159265358979f);
    return num > 0.f ? 90.f : (num < 0.f ? -90.f : 0.f);
}
__global__ void k2_adcig(Ex2 ex, G2 g, int nang, float dang, float* __restrict__ pp, float* __restrict__ ps,
                         float* __restrict__ ppc, float* __restrict__ psc, float* __restrict__ gam) {
    const long mn = (long)g.nx * g.nz;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < mn; q += (long)gridDim.x * blockDim.x) {
        const int ix = (int)(q / g.nz), iz = (int)(q % g.nz);
        const int i = ix + g.bl, k = iz + g.bu;
        const long n = (long)i * g.pz + k;
        float gm = 0.f;
        if (ex.time[n] >= 0) {
            float sxz = 0.f, szz = 0.f, rxz = 0.f, rzz = 0.f;
            for (int di = -2; di <= 2; ++di)
                for (int dk = -2; dk <= 2; ++dk) {
                    const int ii = min(max(i + di, 0), g.nnx - 1), kk = min(max(k + dk, 0), g.nnz - 1);
                    const long m = (long)ii * g.pz + kk;
                    const float t = ex.tp[m], sx = -t * ex.vxp[m], sz = -t * ex.vzp[m];
                    sxz += sx * sz;
                    szz += sz * sz;
                    const float rx = ex.rvxp[m], rz = ex.rvzp[m];
                    rxz += rx * rz;
                    rzz += rz * rz;
                }
            const float ts = deg_atan_ratio(sxz, szz), tr = deg_atan_ratio(rxz, rzz);
            gm = 0.5f * (ts - tr);
            const int b = (int)floorf((gm + 0.5f * nang * dang) / dang);
            if (b >= 0 && b < nang) {
                if (pp) pp[(long)b * mn + q] = ex.img_pp[n];
                if (ps) ps[(long)b * mn + q] = ex.img_ps[n];
            }
        }
        float cp = 0.f, cs = 0.f;
        if (ex.time[n] >= 0) {
            const float o = gm * (2.f * 3.14159265358979f / 180.f), co = cosf(o);
            cp = (co > 0.f ? ex.img_pp[n] : -ex.img_pp[n]) / (fabsf(co) + 0.03f);
            cs = ex.img_ps[n] / (fabsf(sinf(o)) + 0.1f);
        }
        if (ppc) ppc[q] = cp;
        if (psc) psc[q] = cs;
        if (gam) gam[q] = gm;
    }
}
__global__ void k2_fill(float* __restrict__ p, long n, float v) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = v;
}
__global__ void k2_filli(int* __restrict__ p, long n, int v) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = v;
}

}  // namespace erm2
