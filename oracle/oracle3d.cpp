/*
 * oracle3d.cpp - CPU restatement of the reference's 3D elastic RTM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle (FP64
 * instantiation) and the CPU baseline (FP32 instantiation, OpenMP over z).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the timed CPU baseline; the product
 * library (liberm.so) never links or calls it.
 *
 * It follows, step for step, the reference files:
 *   step order of the three time loops  3D_elastic_RTM/3D_forward_or_back_together.cu:9-155,161-290,296-453
 *   staggered operators / update form   2D_elastic_RTM/viscoelastic_equation.cu:137-176,294-323
 *                                       (elastic limit: tao=strain so memory terms vanish, :43-56)
 *   vector P (tp, vxp/vyp/vzp)          3D_forward_or_back_together.cu:83,118 (argument lists);
 *                                       tp update as viscoelastic_equation.cu:763
 *   damping profile 1-cos               3D_elastic_RTM/3D_zzzzz:261-369
 *   model padding                       3D_elastic_RTM/3D_zzzzz:141-259
 *   Ricker                              3D_elastic_RTM/zzzzz:136-157 (make_ricker_new)
 *   FD coefficients (radius 6)          3D_elastic_RTM/zzzzz:202-219
 *   vs, density from vp                 3D_elastic_RTM/zzzzz:1261-1287
 *   excitation amplitude/time           2D_elastic_RTM/elastic_2D_kernel_3.cu:30-48,71-89
 *   excitation-amplitude imaging        2D_elastic_RTM/elastic_2D_kernel_3.cu:542-580
 * C-PML (boundary == ERM_BOUNDARY_CPML) has no reference counterpart: it is
 * the standard convolutional PML (Roden & Gedney 2000, Komatitsch & Martin
 * 2007) with kappa = 1, pinned by tests/test_cpml.py (exact no-op at Rc = 1,
 * reflection against an enlarged domain).
 * The 3D kernel bodies are absent from the reference; the 3D staggering
 * (vy at (i,j-1/2,k), sxy at (i-1/2,j-1/2,k), syz at (i,j-1/2,k+1/2)) extends
 * the 2D positions of viscoelastic_equation.cu by symmetry.
 *
 * How it is pinned (tests/test_oracle3d.py): coefficient sums and polynomial
 * exactness, exact decay under uniform damping, x/y mirror symmetry of an
 * explosive source, y<->x transposition symmetry, P arrival times against
 * r/vp, S-energy leakage of the P/S split, excitation time against the
 * travel time, conservation of the discrete leapfrog energy without damping
 * or source, and the FP32 instantiation against the FP64 one; the
 * cross-correlation shot (oracle3d_xcorr_shot, full source history in memory)
 * by a flat reflector imaged at its depth.  The direct-wave mute is restated
 * in oracle3d.py (mute_direct) and pinned there.
 */
#include "../include/erm.h"
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Reference radius-6 optimized first-derivative coefficients, zzzzz:205-210. */
const double kCoeR6[7] = {0.0, 1.2213363647, -0.096931457519, 0.017447662353,
                          -0.0029672895159, 0.00035900539822, -0.00002184781161};
/* Radius-4 staggered Taylor coefficients (exact for polynomials of degree <= 8). */
const double kCoeR4[5] = {0.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0};

/* make_ricker_new, zzzzz:136-157: dt in seconds, PI of cwp.h. */
template <typename T>
std::vector<T> ricker(double freq, double dt_s) {
    const double PI = 3.141592653589793;
    if (freq == 0.0) freq = 30.0;
    if (dt_s == 0.0) dt_s = 0.004;
    T B = (T)(std::sqrt(6.0) / (PI * freq));
    int N = (int)std::ceil(1.35 * (double)B / dt_s);
    std::vector<T> w(2 * N + 1);
    w[N] = (T)1.0;
    for (int i = 1; i <= N; ++i) {
        T t = (T)dt_s * (T)i;
        T u = (T)(2.0 * std::sqrt(6.0)) * t / B;
        w[N + i] = w[N - i] = (T)0.5 * ((T)2.0 - u * u) * std::exp(-u * u / (T)4.0);
    }
    return w;
}

/* One face of coe_attenuation_3d_new2 (3D_zzzzz:284-313): d[i] for i in [0,b),
 * `rising` = profile grows with i (down/right/back faces). pai = 3.1415926 as the reference. */
void face_profile(int b, double c, bool rising, std::vector<double>& d) {
    const double pai = 3.1415926;
    d.assign(b, 0.0);
    for (int i = 0; i < b; ++i) {
        double s = rising ? (double)i / (b - 1) : (double)(b - i - 1) / (b - 1);
        d[i] = c * (1.0 - std::cos(pai / 2 * s));
    }
}

template <typename T>
struct Oracle3D {
    erm_params p;
    int R = 6;
    int nnx = 0, nny = 0, nnz = 0;
    size_t N = 0;
    T c[7] = {0};
    T cx = 0, cy = 0, cz = 0, dtr = 0;
    std::vector<T> f[13];                 /* vx vy vz txx tyy tzz txy txz tyz tp vxp vyp vzp */
    std::vector<T> M[2], MU[2], B[2];     /* rho*vp^2, rho*vs^2, 1/rho; [0]=true [1]=smoothed */
    bool have_model[2] = {false, false};
    std::vector<T> ax, ay, az;            /* damping per axis; att(i,j,k) = ax[i]+ay[j]+az[k] */
    std::vector<T> wavelet;
    std::vector<int> ex_time;
    std::vector<T> ex_amp, ex_tp, ex_vxp, ex_vyp, ex_vzp;
    std::vector<T> img_tp, img_pp, img_ps;
    std::vector<T> poyn[6];               /* receiver Poynting rp, rs at the imaging step (p.poynting) */
    std::vector<T> rec;                   /* [3][lt][ry][rx] */
    /* CPML (no reference counterpart): per axis a, b at integer ("full") and
     * staggered ("half": x-1/2, y-1/2, z+1/2) positions, and 21 memory
     * variables kept on the whole grid (zero outside the absorbing layers). */
    bool cpml = false;
    std::vector<T> ca[3][2], cb[3][2];    /* [axis][0 full, 1 half] */
    std::vector<T> psi[21];
    double vmax_loaded = 0.0;
    int sx = 0, sy = 0, sz = 0;           /* padded-grid source position */

    size_t id(int i, int j, int k) const { return ((size_t)k * nny + j) * nnx + i; }

    int init(const erm_params* pp) {
        p = *pp;
        if (p.fd_order == 12) R = 6;
        else if (p.fd_order == 8) R = 4;
        else return ERM_E_PARAM;
        if (p.boundary != ERM_BOUNDARY_DAMP && p.boundary != ERM_BOUNDARY_CPML) return ERM_E_PARAM;
        nnx = p.nx + p.bl + p.br;
        nny = p.ny + p.bf + p.bb;
        nnz = p.nz + p.bu + p.bd;
        if (nnx < 2 * R + 1 || nny < 2 * R + 1 || nnz < 2 * R + 1) return ERM_E_SIZE;
        N = (size_t)nnx * nny * nnz;
        for (int m = 0; m <= R; ++m) c[m] = (T)(R == 6 ? kCoeR6[m] : kCoeR4[m]);
        dtr = (T)(p.dt / 1000.0);
        cx = (T)((p.dt / 1000.0) / p.dx);
        cy = (T)((p.dt / 1000.0) / p.dy);
        cz = (T)((p.dt / 1000.0) / p.dz);
        for (auto& v : f) v.assign(N, (T)0);
        /* damping, 3D_zzzzz:261-369 (sum over the six faces) */
        std::vector<double> du, dd, dl, dr, df, db;
        double cA = p.boundary == ERM_BOUNDARY_CPML ? 0.0 : p.coe_attenuation;  /* CPML replaces damping */
        face_profile(p.bu, cA, false, du);
        face_profile(p.bd, cA, true, dd);
        face_profile(p.bl, cA, false, dl);
        face_profile(p.br, cA, true, dr);
        face_profile(p.bf, cA, false, df);
        face_profile(p.bb, cA, true, db);
        ax.assign(nnx, 0); ay.assign(nny, 0); az.assign(nnz, 0);
        for (int i = 0; i < p.bl; ++i) ax[i] += (T)dl[i];
        for (int i = 0; i < p.br; ++i) ax[p.bl + p.nx + i] += (T)dr[i];
        for (int j = 0; j < p.bf; ++j) ay[j] += (T)df[j];
        for (int j = 0; j < p.bb; ++j) ay[p.bf + p.ny + j] += (T)db[j];
        for (int k = 0; k < p.bu; ++k) az[k] += (T)du[k];
        for (int k = 0; k < p.bd; ++k) az[p.bu + p.nz + k] += (T)dd[k];
        wavelet = ricker<T>(p.freq, p.dt / 1000.0);
        ex_time.assign(N, -1);
        ex_amp.assign(N, 0); ex_tp.assign(N, 0); ex_vxp.assign(N, 0); ex_vyp.assign(N, 0); ex_vzp.assign(N, 0);
        img_tp.assign(N, 0); img_pp.assign(N, 0); img_ps.assign(N, 0);
        for (auto& v : poyn) v.assign(p.poynting == 1 ? N : 0, 0);
        rec.assign((size_t)3 * p.lt * p.receiver_num_y * p.receiver_num_x, 0);
        return set_shot(0, 0, 0);
    }

    int set_shot(int isx, int isy, int isz) {
        int sxr = p.shot_start_x + isx * p.shot_interval_x;
        int syr = p.shot_start_y + isy * p.shot_interval_y;
        int szr = p.shot_start_z + isz * p.shot_interval_z;
        sx = p.bl + sxr; sy = p.bf + syr; sz = p.bu + szr;
        if (sx < R || sx >= nnx - R || sy < R || sy >= nny - R || sz < R || sz >= nnz - R) return ERM_E_PARAM;
        return ERM_OK;
    }

    /* add_pml_layers_v_h (3D_zzzzz:141-259): edge replication == clamping
     * the padded index into the model region. */
    void pad(const float* src, std::vector<double>& dst) const {
        dst.resize(N);
        for (int k = 0; k < nnz; ++k) {
            int kz = std::min(std::max(k - p.bu, 0), p.nz - 1);
            for (int j = 0; j < nny; ++j) {
                int jy = std::min(std::max(j - p.bf, 0), p.ny - 1);
                for (int i = 0; i < nnx; ++i) {
                    int ix = std::min(std::max(i - p.bl, 0), p.nx - 1);
                    dst[id(i, j, k)] = src[((size_t)ix * p.ny + jy) * p.nz + kz];
                }
            }
        }
    }

    int load_model(int which, const float* vp, const float* vs, const float* rho) {
        if (which < 0 || which > 1 || !vp) return ERM_E_PARAM;
        std::vector<double> pvp, pvs, prho;
        pad(vp, pvp);
        if (vs) pad(vs, pvs);
        else { pvs.resize(N); for (size_t n = 0; n < N; ++n) pvs[n] = (float)(pvp[n] / std::sqrt(3.0)); }
        if (rho) pad(rho, prho);
        else { prho.resize(N); for (size_t n = 0; n < N; ++n) prho[n] = (float)(1000 * 0.23 * std::sqrt(std::sqrt(pvp[n]))); }
        M[which].resize(N); MU[which].resize(N); B[which].resize(N);
        for (size_t n = 0; n < N; ++n) {
            T r = (T)prho[n], a = (T)pvp[n], b = (T)pvs[n];
            M[which][n] = r * a * a;
            MU[which][n] = r * b * b;
            B[which][n] = (T)1 / r;
        }
        have_model[which] = true;
        double vm = 0;
        for (size_t n = 0; n < N; ++n) vm = std::max(vm, pvp[n]);
        vmax_loaded = std::max(vmax_loaded, vm);
        if (p.boundary == ERM_BOUNDARY_CPML) build_cpml();
        return ERM_OK;
    }

    /* C-PML (Roden & Gedney 2000; Komatitsch & Martin 2007) with kappa = 1:
     * psi' = b psi + a e,  e' = e + psi',  b = exp(-(d+alpha) dt),
     * a = d (b-1)/(d+alpha);  d = d0 r^2, d0 = -3 vmax ln(Rc) / (2 L),
     * alpha = pi f0 (1-r), r = distance into the layer / L, L = width - R. */
    void build_cpml() {
        cpml = true;
        const double vmax = p.cpml_vmax > 0 ? p.cpml_vmax : vmax_loaded;
        const double Rc = p.cpml_r > 0 ? p.cpml_r : 1e-3;
        const double dts = p.dt / 1000.0;
        const int n3[3] = {nnx, nny, nnz};
        const int lo[3] = {p.bl, p.bf, p.bu}, hi[3] = {p.br, p.bb, p.bd};
        const double h[3] = {p.dx, p.dy, p.dz};
        const double half_shift[3] = {-0.5, -0.5, +0.5};
        for (int a = 0; a < 3; ++a)
            for (int f = 0; f < 2; ++f) {
                ca[a][f].assign(n3[a], 0);
                cb[a][f].assign(n3[a], 1);
                for (int i = 0; i < n3[a]; ++i) {
                    double x = i + (f ? half_shift[a] : 0.0);
                    /* the layer spans the updated cells only: the outer R
                     * cells are never updated, so L = width - R */
                    double s = 0, L = 0;
                    if (lo[a] > 0 && x < lo[a]) { s = lo[a] - x; L = std::max(lo[a] - R, 1); }
                    else if (hi[a] > 0 && x > n3[a] - hi[a] - 1) { s = x - (n3[a] - hi[a] - 1); L = std::max(hi[a] - R, 1); }
                    if (L <= 0) continue;
                    double r = s / L;
                    double d0 = -3.0 * vmax * std::log(Rc) / (2.0 * L * h[a]);
                    double d = d0 * r * r;
                    double al = M_PI * p.freq * (1.0 - std::min(r, 1.0));
                    double bb = std::exp(-(d + al) * dts);
                    cb[a][f][i] = (T)bb;
                    ca[a][f][i] = (T)(d > 0 ? d * (bb - 1.0) / (d + al) : 0.0);
                }
            }
        for (auto& v : psi) v.assign(N, (T)0);
    }
    /* one memory-variable update: e += (psi = b psi + a e) */
    inline T pml(int q, size_t n, int axis, int half, int idx, T e) {
        const T bq = cb[axis][half][idx], aq = ca[axis][half][idx];
        if (aq == (T)0) return e;
        T v = bq * psi[q][n] + aq * e;
        psi[q][n] = v;
        return e + v;
    }

    /* D+ f[i] = sum c_m (f[i+m] - f[i-m+1])  (centred at i+1/2),
     * D- f[i] = sum c_m (f[i+m-1] - f[i-m])  (centred at i-1/2);
     * viscoelastic_equation.cu:137-149. */
    inline T dp(const T* q, ptrdiff_t s) const {
        T a = 0;
        for (int m = 1; m <= R; ++m) a += c[m] * (q[m * s] - q[-(m - 1) * s]);
        return a;
    }
    inline T dm(const T* q, ptrdiff_t s) const {
        T a = 0;
        for (int m = 1; m <= R; ++m) a += c[m] * (q[(m - 1) * s] - q[-m * s]);
        return a;
    }

    void add_source(int it) {  /* add_source_3D, 3D_forward_or_back_together.cu:36-38 */
        if (it < 0 || it >= (int)wavelet.size()) return;
        size_t s = id(sx, sy, sz);
        f[3][s] += wavelet[it];
        f[4][s] += wavelet[it];
        f[5][s] += wavelet[it];
    }

    /* fwd_txxzzxzpp_3D: stress and P-stress from v (launch :83). */
    void stress(int which) {
        const T* vx = f[0].data(); const T* vy = f[1].data(); const T* vz = f[2].data();
        T* txx = f[3].data(); T* tyy = f[4].data(); T* tzz = f[5].data();
        T* txy = f[6].data(); T* txz = f[7].data(); T* tyz = f[8].data(); T* tp = f[9].data();
        const T* Mm = M[which].data(); const T* Um = MU[which].data();
        const ptrdiff_t sX = 1, sY = nnx, sZ = (ptrdiff_t)nnx * nny;
        const T h = dtr / (T)2;
#pragma omp parallel for schedule(static)
        for (int k = R; k < nnz - R; ++k)
            for (int j = R; j < nny - R; ++j)
                for (int i = R; i < nnx - R; ++i) {
                    size_t n = id(i, j, k);
                    T exx = cx * dp(vx + n, sX), eyy = cy * dp(vy + n, sY), ezz = cz * dm(vz + n, sZ);
                    T e_yx = cy * dm(vx + n, sY), e_xy = cx * dm(vy + n, sX);
                    T e_zx = cz * dp(vx + n, sZ), e_xz = cx * dm(vz + n, sX);
                    T e_zy = cz * dp(vy + n, sZ), e_yz = cy * dm(vz + n, sY);
                    if (cpml) {
                        exx = pml(0, n, 0, 0, i, exx); e_xy = pml(1, n, 0, 1, i, e_xy); e_xz = pml(2, n, 0, 1, i, e_xz);
                        eyy = pml(3, n, 1, 0, j, eyy); e_yx = pml(4, n, 1, 1, j, e_yx); e_yz = pml(5, n, 1, 1, j, e_yz);
                        ezz = pml(6, n, 2, 0, k, ezz); e_zx = pml(7, n, 2, 1, k, e_zx); e_zy = pml(8, n, 2, 1, k, e_zy);
                    }
                    T exy = e_yx + e_xy, exz = e_zx + e_xz, eyz = e_zy + e_yz;
                    T d = ax[i] + ay[j] + az[k];
                    T a = (T)1 / ((T)1 + d * h), b = (T)1 - d * h;
                    T Mc = Mm[n], mu = Um[n], lam = Mc - (T)2 * mu;
                    txx[n] = a * (b * txx[n] + Mc * exx + lam * (eyy + ezz));
                    tyy[n] = a * (b * tyy[n] + Mc * eyy + lam * (exx + ezz));
                    tzz[n] = a * (b * tzz[n] + Mc * ezz + lam * (exx + eyy));
                    tp[n] = a * (b * tp[n] + Mc * (exx + eyy + ezz));
                    txy[n] = a * (b * txy[n] + mu * exy);
                    txz[n] = a * (b * txz[n] + mu * exz);
                    tyz[n] = a * (b * tyz[n] + mu * eyz);
                }
    }

    /* fwd_vx_3D/fwd_vy_3D/fwd_vz_3D (:100-104) + fwd_vxp_vzp_3D (:118),
     * followed by excitation (:282, mode 1) or imaging (:442-446, mode 2). */
    void velocity(int which, int mode, int it) {
        T* vx = f[0].data(); T* vy = f[1].data(); T* vz = f[2].data();
        const T* txx = f[3].data(); const T* tyy = f[4].data(); const T* tzz = f[5].data();
        const T* txy = f[6].data(); const T* txz = f[7].data(); const T* tyz = f[8].data();
        const T* tp = f[9].data();
        T* vxp = f[10].data(); T* vyp = f[11].data(); T* vzp = f[12].data();
        const T* Bm = B[which].data();
        const ptrdiff_t sX = 1, sY = nnx, sZ = (ptrdiff_t)nnx * nny;
        const T h = dtr / (T)2;
#pragma omp parallel for schedule(static)
        for (int k = R; k < nnz - R; ++k)
            for (int j = R; j < nny - R; ++j)
                for (int i = R; i < nnx - R; ++i) {
                    size_t n = id(i, j, k);
                    T d = ax[i] + ay[j] + az[k];
                    T a = (T)1 / ((T)1 + d * h), b = (T)1 - d * h;
                    T bo = Bm[n];
                    T gxx = cx * dm(txx + n, sX), gyx = cy * dp(txy + n, sY), gzx = cz * dm(txz + n, sZ);
                    T gxy = cx * dp(txy + n, sX), gyy = cy * dm(tyy + n, sY), gzy = cz * dm(tyz + n, sZ);
                    T gxz = cx * dp(txz + n, sX), gyz = cy * dp(tyz + n, sY), gzz = cz * dp(tzz + n, sZ);
                    T px = cx * dm(tp + n, sX), py = cy * dm(tp + n, sY), pz = cz * dp(tp + n, sZ);
                    if (cpml) {
                        gxx = pml(9, n, 0, 1, i, gxx); gxy = pml(10, n, 0, 0, i, gxy);
                        gxz = pml(11, n, 0, 0, i, gxz); px = pml(12, n, 0, 1, i, px);
                        gyx = pml(13, n, 1, 0, j, gyx); gyy = pml(14, n, 1, 1, j, gyy);
                        gyz = pml(15, n, 1, 0, j, gyz); py = pml(16, n, 1, 1, j, py);
                        gzx = pml(17, n, 2, 0, k, gzx); gzy = pml(18, n, 2, 0, k, gzy);
                        gzz = pml(19, n, 2, 1, k, gzz); pz = pml(20, n, 2, 1, k, pz);
                    }
                    T fx = gxx + gyx + gzx;
                    T fy = gxy + gyy + gzy;
                    T fz = gxz + gyz + gzz;
                    T nvx = a * (b * vx[n] + bo * fx);
                    T nvy = a * (b * vy[n] + bo * fy);
                    T nvz = a * (b * vz[n] + bo * fz);
                    T nxp = a * (b * vxp[n] + bo * px);
                    T nyp = a * (b * vyp[n] + bo * py);
                    T nzp = a * (b * vzp[n] + bo * pz);
                    vx[n] = nvx; vy[n] = nvy; vz[n] = nvz;
                    vxp[n] = nxp; vyp[n] = nyp; vzp[n] = nzp;
                    if (mode == 1) excite_cell(n, it, tp[n], nxp, nyp, nzp);
                    else if (mode == 2) image_cell(n, it, tp[n], nvx, nvy, nvz, nxp, nyp, nzp);
                }
    }

    /* caculate_ex_amp_time_new (elastic_2D_kernel_3.cu:30-48) in 3D, the
     * cuda_cal_excitation_amp_time_new launch (:282): amplitude |vP|, strict
     * '<' so the first maximum wins; tp and vP of that step are kept. */
    void excite_cell(size_t n, int it, T tpv, T nxp, T nyp, T nzp) {
        T amp = std::sqrt(nxp * nxp + nyp * nyp + nzp * nzp);
        if (ex_amp[n] < amp) {
            ex_time[n] = it; ex_amp[n] = amp; ex_tp[n] = tpv;
            ex_vxp[n] = nxp; ex_vyp[n] = nyp; ex_vzp[n] = nzp;
        }
    }
    /* imaging_correlation_ex + imaging_inner_product_ex (:442-446; 2D bodies
     * elastic_2D_kernel_3.cu:542-580) at the cell's excitation step:
     *   I_tp = tp_r / ex_tp,  I_PP = ex_vP . vP_r / |ex_vP|^2,
     *   I_PS = ex_vP . (v_r - vP_r) / |ex_vP|^2;
     * a zero denominator images 0.  The reference passes the stabilisers
     * amp_max = tp_max = 0 (globals, 3D_elastic_modeling_parameter:139, never
     * set in the shipped files), so the plain quotients are its result. */
    void image_cell(size_t n, int it, T tpv, T nvx, T nvy, T nvz, T nxp, T nyp, T nzp) {
        if (ex_time[n] != it) return;
        T etp = ex_tp[n];
        img_tp[n] = etp != (T)0 ? tpv / etp : (T)0;
        T ex = ex_vxp[n], ey = ex_vyp[n], ez = ex_vzp[n];
        T den = ex * ex + ey * ey + ez * ez;
        T pp = ex * nxp + ey * nyp + ez * nzp;
        T ps = ex * (nvx - nxp) + ey * (nvy - nyp) + ez * (nvz - nzp);
        img_pp[n] = den > (T)0 ? pp / den : (T)0;
        img_ps[n] = den > (T)0 ? ps / den : (T)0;
        if (!poyn[0].empty()) poynting_cell(n, tpv, nvx, nvy, nvz, nxp, nyp, nzp);
    }
    /* cuda_cal_receiver_poyn_3D (call commented out at :432, body absent) as
     * the 3D form of sum_poynting (elastic_2D_kernel_2.cu:8-51) at the cell's
     * imaging step: rp = -tp vP, rs = -(sigma - tp I) vS with vS = v - vP,
     * every product at the cell's own index. */
    void poynting_cell(size_t n, T tpv, T nvx, T nvy, T nvz, T nxp, T nyp, T nzp) {
        const T sx = nvx - nxp, sy = nvy - nyp, sz = nvz - nzp;
        const T xx = f[3][n] - tpv, yy = f[4][n] - tpv, zz = f[5][n] - tpv;
        const T xy = f[6][n], xz = f[7][n], yz = f[8][n];
        poyn[0][n] = -tpv * nxp; poyn[1][n] = -tpv * nyp; poyn[2][n] = -tpv * nzp;
        poyn[3][n] = -(xx * sx + xy * sy + xz * sz);
        poyn[4][n] = -(xy * sx + yy * sy + yz * sz);
        poyn[5][n] = -(xz * sx + yz * sy + zz * sz);
    }
    /* The imaging condition alone, applied to the current fields at step it
     * (no propagation): pins the condition on prescribed receiver fields. */
    void image_fields(int it) {
        for (int k = R; k < nnz - R; ++k)
            for (int j = R; j < nny - R; ++j)
                for (int i = R; i < nnx - R; ++i) {
                    size_t n = id(i, j, k);
                    image_cell(n, it, f[9][n], f[0][n], f[1][n], f[2][n], f[10][n], f[11][n], f[12][n]);
                }
    }

    size_t rec_index(int c3, int it, int iry, int irx) const {
        return (((size_t)c3 * p.lt + it) * p.receiver_num_y + iry) * p.receiver_num_x + irx;
    }
    template <typename F>
    void for_receivers(F&& fn) const {
        for (int iry = 0; iry < p.receiver_num_y; ++iry)
            for (int irx = 0; irx < p.receiver_num_x; ++irx) {
                int i = p.bl + p.receiver_start_x + irx * p.receiver_interval_x;
                int j = p.bf + p.receiver_start_y + iry * p.receiver_interval_y;
                int k = p.bu + p.receiver_start_z;
                fn(iry, irx, id(i, j, k));
            }
    }

    int check_receivers() const {
        int bad = 0;
        for_receivers([&](int, int, size_t n) { if (n >= N) bad = 1; });
        int i1 = p.bl + p.receiver_start_x + (p.receiver_num_x - 1) * p.receiver_interval_x;
        int j1 = p.bf + p.receiver_start_y + (p.receiver_num_y - 1) * p.receiver_interval_y;
        int k0 = p.bu + p.receiver_start_z;
        if (p.bl + p.receiver_start_x < 0 || i1 >= nnx || p.bf + p.receiver_start_y < 0 || j1 >= nny ||
            k0 < 0 || k0 >= nnz)
            bad = 1;
        return bad ? ERM_E_PARAM : ERM_OK;
    }

    int step(int pass, int it_begin, int nsteps) {
        int which = pass == ERM_PASS_MODEL ? 0 : 1;
        if (!have_model[which]) return ERM_E_STATE;
        if (check_receivers() != ERM_OK) return ERM_E_PARAM;
        for (int s = 0; s < nsteps; ++s) {
            if (pass == ERM_PASS_RECEIVER) {
                int it = it_begin - s;
                if (it < 0 || it >= p.lt) return ERM_E_PARAM;
                /* write_or_add_shot_3D_surface_three, mode add (:328) into v at level it */
                for_receivers([&](int iry, int irx, size_t n) {
                    f[0][n] += rec[rec_index(0, it, iry, irx)];
                    f[1][n] += rec[rec_index(1, it, iry, irx)];
                    f[2][n] += rec[rec_index(2, it, iry, irx)];
                });
                stress(which);
                velocity(which, 2, it);
            } else {
                int it = it_begin + s;
                add_source(it);
                stress(which);
                velocity(which, pass == ERM_PASS_SOURCE ? 1 : 0, it);
                if (pass == ERM_PASS_MODEL && it < p.lt) {
                    /* write_or_add_shot_3D_surface_three, mode 0 (:127) */
                    for_receivers([&](int iry, int irx, size_t n) {
                        rec[rec_index(0, it, iry, irx)] = f[0][n];
                        rec[rec_index(1, it, iry, irx)] = f[1][n];
                        rec[rec_index(2, it, iry, irx)] = f[2][n];
                    });
                }
            }
        }
        return ERM_OK;
    }

    void zero_wavefield() {
        for (auto& v : f) std::fill(v.begin(), v.end(), (T)0);
        for (auto& v : psi) std::fill(v.begin(), v.end(), (T)0);
    }
    void zero_excitation() {
        std::fill(ex_time.begin(), ex_time.end(), -1);
        for (auto* v : {&ex_amp, &ex_tp, &ex_vxp, &ex_vyp, &ex_vzp}) std::fill(v->begin(), v->end(), (T)0);
        for (auto& v : poyn) std::fill(v.begin(), v.end(), (T)0);
    }
    void zero_images() {
        for (auto* v : {&img_tp, &img_pp, &img_ps}) std::fill(v->begin(), v->end(), (T)0);
        for (auto& v : poyn) std::fill(v.begin(), v.end(), (T)0);
    }

    /* P/S decomposition by divergence and curl (the 3D form of decom,
     * elastic_2D_kernel_2.cu:621-666, with the y positions of the 3D grid):
     * q = 0 div at (i,j,k); 1 curl_x at the syz node; 2 curl_y at the sxz node;
     * 3 curl_z at the sxy node; interior cells, 0 elsewhere. */
    void decom(int q, double* dst) const {
        const T* vx = f[0].data(); const T* vy = f[1].data(); const T* vz = f[2].data();
        const ptrdiff_t sX = 1, sY = nnx, sZ = (ptrdiff_t)nnx * nny;
        const T ix = (T)(1.0 / p.dx), iy = (T)(1.0 / p.dy), iz = (T)(1.0 / p.dz);
        for (size_t n = 0; n < N; ++n) dst[n] = 0.0;
        for (int k = R; k < nnz - R; ++k)
            for (int j = R; j < nny - R; ++j)
                for (int i = R; i < nnx - R; ++i) {
                    size_t n = id(i, j, k);
                    T v;
                    if (q == 0) v = ix * dp(vx + n, sX) + iy * dp(vy + n, sY) + iz * dm(vz + n, sZ);
                    else if (q == 1) v = iy * dm(vz + n, sY) - iz * dp(vy + n, sZ);
                    else if (q == 2) v = iz * dp(vx + n, sZ) - ix * dm(vz + n, sX);
                    else v = ix * dm(vy + n, sX) - iy * dm(vx + n, sY);
                    dst[n] = (double)v;
                }
    }

    int get_field(int fid, double* dst) const {
        if (fid >= 0 && fid < 13) { for (size_t n = 0; n < N; ++n) dst[n] = (double)f[fid][n]; return 0; }
        if (fid >= ERM_F_DIV && fid <= ERM_F_CURLZ) { decom(fid - ERM_F_DIV, dst); return 0; }
        if (fid >= ERM_F_VXS && fid <= ERM_F_VZS) {
            int c3 = fid - ERM_F_VXS;
            for (size_t n = 0; n < N; ++n) dst[n] = (double)(f[c3][n] - f[10 + c3][n]);
            return 0;
        }
        if (fid >= ERM_F_POYN_SX && fid <= ERM_F_POYN_SZ) {  /* source P at the excitation step */
            const std::vector<T>& v = fid == ERM_F_POYN_SX ? ex_vxp : fid == ERM_F_POYN_SY ? ex_vyp : ex_vzp;
            for (size_t n = 0; n < N; ++n) dst[n] = (double)(-ex_tp[n] * v[n]);
            return 0;
        }
        if (fid >= ERM_F_POYN_RPX && fid <= ERM_F_POYN_RSZ) {
            const std::vector<T>& v = poyn[fid - ERM_F_POYN_RPX];
            if (v.empty()) return ERM_E_STATE;
            for (size_t n = 0; n < N; ++n) dst[n] = (double)v[n];
            return 0;
        }
        const std::vector<T>* src = nullptr;
        switch (fid) {
            case ERM_F_EX_TIME: for (size_t n = 0; n < N; ++n) dst[n] = ex_time[n]; return 0;
            case ERM_F_EX_AMP: src = &ex_amp; break;
            case ERM_F_EX_TP: src = &ex_tp; break;
            case ERM_F_EX_VXP: src = &ex_vxp; break;
            case ERM_F_EX_VYP: src = &ex_vyp; break;
            case ERM_F_EX_VZP: src = &ex_vzp; break;
            case ERM_F_IMG_TP: src = &img_tp; break;
            case ERM_F_IMG_PP: src = &img_pp; break;
            case ERM_F_IMG_PS: src = &img_ps; break;
            default: return ERM_E_PARAM;
        }
        for (size_t n = 0; n < N; ++n) dst[n] = (double)(*src)[n];
        return 0;
    }
    int set_field(int fid, const double* src) {
        if (fid >= 0 && fid < 13) { for (size_t n = 0; n < N; ++n) f[fid][n] = (T)src[n]; return 0; }
        return ERM_E_PARAM;
    }
};

struct Handle {
    int prec;
    Oracle3D<double>* d = nullptr;
    Oracle3D<float>* s = nullptr;
};

template <typename Fn>
int dispatch(void* h, Fn&& fn) {
    if (!h) return ERM_E_PARAM;
    Handle* H = (Handle*)h;
    if (H->prec == 64) return fn(*H->d);
    return fn(*H->s);
}

}  // namespace

extern "C" {

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int oracle_coefficients(int radius, double* c) {
    if (radius == 6) { for (int m = 0; m <= 6; ++m) c[m] = kCoeR6[m]; return 0; }
    if (radius == 4) { for (int m = 0; m <= 4; ++m) c[m] = kCoeR4[m]; return 0; }
    return ERM_E_PARAM;
}

int oracle_ricker(double freq, double dt_ms, double* w, int cap) {
    std::vector<double> r = ricker<double>(freq, dt_ms / 1000.0);
    if (w) for (int i = 0; i < (int)r.size() && i < cap; ++i) w[i] = r[i];
    return (int)r.size();
}

int oracle_damping_axis(const erm_params* p, int axis, double* out) {
    /* per-axis damping sum used by oracle3d (x: nnx, y: nny, z: nnz entries) */
    Oracle3D<double> o;
    int rc = o.init(p);
    if (rc) return rc;
    const std::vector<double>& v = axis == 0 ? o.ax : axis == 1 ? o.ay : o.az;
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    return (int)v.size();
}

void* oracle3d_create(const erm_params* p, int prec) {
    Handle* H = new Handle();
    H->prec = prec == 32 ? 32 : 64;
    int rc;
    if (H->prec == 64) { H->d = new Oracle3D<double>(); rc = H->d->init(p); }
    else { H->s = new Oracle3D<float>(); rc = H->s->init(p); }
    if (rc) { delete H->d; delete H->s; delete H; return nullptr; }
    return H;
}

void oracle3d_destroy(void* h) {
    if (!h) return;
    Handle* H = (Handle*)h;
    delete H->d; delete H->s; delete H;
}

int oracle3d_load_model(void* h, int which, const float* vp, const float* vs, const float* rho) {
    return dispatch(h, [&](auto& o) { return o.load_model(which, vp, vs, rho); });
}
int oracle3d_set_shot(void* h, int isx, int isy, int isz) {
    return dispatch(h, [&](auto& o) { return o.set_shot(isx, isy, isz); });
}
int oracle3d_zero(void* h, int what) {
    return dispatch(h, [&](auto& o) {
        if (what & 1) o.zero_wavefield();
        if (what & 2) o.zero_excitation();
        if (what & 4) o.zero_images();
        return 0;
    });
}
int oracle3d_step(void* h, int pass, int it_begin, int nsteps) {
    return dispatch(h, [&](auto& o) { return o.step(pass, it_begin, nsteps); });
}
int oracle3d_get_field(void* h, int fid, double* dst) {
    return dispatch(h, [&](auto& o) { return o.get_field(fid, dst); });
}
int oracle3d_set_field(void* h, int fid, const double* src) {
    return dispatch(h, [&](auto& o) { return o.set_field(fid, src); });
}
int oracle3d_image(void* h, int it) {
    return dispatch(h, [&](auto& o) { o.image_fields(it); return 0; });
}
int oracle3d_get_record(void* h, double* dst) {
    return dispatch(h, [&](auto& o) {
        for (size_t n = 0; n < o.rec.size(); ++n) dst[n] = (double)o.rec[n];
        return 0;
    });
}
int oracle3d_set_record(void* h, const double* src) {
    return dispatch(h, [&](auto& o) {
        for (size_t n = 0; n < o.rec.size(); ++n) o.rec[n] = (decltype(o.rec[0] + 0))src[n];
        return 0;
    });
}
/* One full shot in the order of the reference driver: modeling in the true
 * model (skipped when use_record != 0), source forward with excitation,
 * receiver backward with imaging. */
int oracle3d_rtm_shot(void* h, int use_record) {
    return dispatch(h, [&](auto& o) {
        int rc;
        if (!use_record) {
            o.zero_wavefield();
            if ((rc = o.step(ERM_PASS_MODEL, 0, o.p.lt))) return rc;
        }
        o.zero_wavefield();
        o.zero_excitation();
        if ((rc = o.step(ERM_PASS_SOURCE, 0, o.p.lt_forward))) return rc;
        o.zero_wavefield();
        o.zero_images();
        return o.step(ERM_PASS_RECEIVER, o.p.lt - 1, o.p.lt);
    });
}

/* Zero-lag cross-correlation imaging with the source wavefield kept in full
 * (the full-history strategy of elastic_2D_kernel_2.cu:793-827, here in
 * memory): source pass storing vP and tp over the model region after every
 * step, then the receiver pass accumulating, after every step it,
 *   I_tp += tp_s(it) tp_r,  I_PP += vP_s(it).vP_r,  I_PS += vP_s(it).(v_r - vP_r).
 * Outputs over the model region in disk order [nx][ny][nz]; requires
 * lt_forward == lt and a record (modelled here when use_record == 0). */
/* decim > 1: the products of the receiver steps it with (it + 1) % decim == 0
 * only, each weighted by decim (the decimated history of
 * erm_forward_smoothed_history) */
int oracle3d_xcorr_shot_decim(void* h, int use_record, int decim, double* itp, double* ipp, double* ips) {
    return dispatch(h, [&](auto& o) {
        using T = typename std::remove_reference<decltype(o.rec[0])>::type;
        const erm_params& p = o.p;
        if (p.lt_forward != p.lt) return ERM_E_PARAM;
        int rc;
        if (!use_record) {
            o.zero_wavefield();
            if ((rc = o.step(ERM_PASS_MODEL, 0, p.lt))) return rc;
        }
        const size_t nm = (size_t)p.nx * p.ny * p.nz;
        auto cell = [&](size_t q) {  // model-region cell q (disk order) -> padded index
            const int iz = (int)(q % p.nz), iy = (int)((q / p.nz) % p.ny), ix = (int)(q / ((size_t)p.nz * p.ny));
            return ((size_t)(iz + p.bu) * o.nny + (iy + p.bf)) * o.nnx + (ix + p.bl);
        };
        std::vector<T> hist((size_t)p.lt * 4 * nm);
        o.zero_wavefield();
        o.zero_excitation();
        for (int it = 0; it < p.lt; ++it) {
            if ((rc = o.step(ERM_PASS_SOURCE, it, 1))) return rc;
            T* h = &hist[(size_t)it * 4 * nm];
            for (size_t q = 0; q < nm; ++q) {
                const size_t n = cell(q);
                h[q] = o.f[9][n]; h[nm + q] = o.f[10][n]; h[2 * nm + q] = o.f[11][n]; h[3 * nm + q] = o.f[12][n];
            }
        }
        for (size_t q = 0; q < nm; ++q) itp[q] = ipp[q] = ips[q] = 0.0;
        o.zero_wavefield();
        o.zero_images();
        for (int it = p.lt - 1; it >= 0; --it) {
            if ((rc = o.step(ERM_PASS_RECEIVER, it, 1))) return rc;
            if ((it + 1) % decim != 0) continue;
            const T* h = &hist[(size_t)it * 4 * nm];
            const double w = decim;
            for (size_t q = 0; q < nm; ++q) {
                const size_t n = cell(q);
                const double sx = h[nm + q], sy = h[2 * nm + q], sz = h[3 * nm + q];
                const double rx = o.f[10][n], ry = o.f[11][n], rz = o.f[12][n];
                itp[q] += w * ((double)h[q] * (double)o.f[9][n]);
                ipp[q] += w * (sx * rx + sy * ry + sz * rz);
                ips[q] += w * (sx * ((double)o.f[0][n] - rx) + sy * ((double)o.f[1][n] - ry) + sz * ((double)o.f[2][n] - rz));
            }
        }
        return 0;
    });
}
int oracle3d_xcorr_shot(void* h, int use_record, double* itp, double* ipp, double* ips) {
    return oracle3d_xcorr_shot_decim(h, use_record, 1, itp, ipp, ips);
}

}  // extern "C"
