/*
 * oracle2d.cpp - CPU restatement of the reference's 2D elastic/viscoelastic
 * kernels and of the LSRTM operators built on them.
 *
 * TEST INFRASTRUCTURE ONLY (FP64 = oracle, FP32 = CPU baseline); loaded by
 * tests/ and bench.py's CPU leg, never by liberm.so.
 *
 * Follows (2D_elastic_RTM/):
 *   SLS parameters, L = 1            viscoelastic_equation.cu:22-58 (cuda_cal_viscoelastic)
 *   stress + memory, 3-param.        viscoelastic_equation.cu:184-325 (tp: :172, :763)
 *   staggered operators              viscoelastic_equation.cu:137-176, adjoint :1923-1967
 *   damping (cubic profile)          3D_elastic_RTM/zzzzz:238-325 (make_attenuation_new)
 *   P/S div/curl decomposition       elastic_2D_kernel_2.cu:563-669 (decom)
 *   excitation, imaging              elastic_2D_kernel_3.cu:30-48, 542-580
 *   Born sources (inversion_para 2)  viscoelastic_equation.cu:2398-2505
 *   adjoint / gradient (structure)   viscoelastic_equation.cu:1827-2295
 * The velocity update (the reference's fwd_vx/fwd_vz kernels are absent) is
 * the 2D form of the stress divergence with the positions of :137-176.
 * L = 3 uses the tau-method (Blanch et al. 1995): tau_eps_l = tau_sig_l (1+tau),
 * tau_sig_l = 1/(2 pi f_l), f_l = freq * {1/4, 1, 4}, tau fixed by Q(2 pi freq) = Q.
 * The adjoint is the exact transpose of the discrete Born operator (pinned
 * by the dot-product test in tests/test_oracle2d.py); adjoint_strips takes the
 * background's E and F from a reversal with ring strips instead of the
 * stored history (save_and_set_wavefiled, viscoelastic_equation.cu:3-21;
 * the reversal of rfwd_..._old, :1298-1435), pinned by equality with the
 * history adjoint and the same dot-product test.
 * The imaging step also keeps the receiver vector P of each cell at its
 * imaging step (r_vxp, r_vzp) for the angle gathers that oracle2d.py restates
 * (pinned by a flat reflector); the source illumination is composed in
 * oracle2d.py from this file's steps (pinned by additivity over shots and its
 * dependence on the background model only).  How each piece is pinned:
 * closed-form derivatives, SLS Q at the reference frequency and over the band,
 * the Q -> infinity limit, conservation of the elastic leapfrog's discrete
 * energy without damping or source, attenuation, Born vs finite differences, the
 * adjoint dot product, a flat reflector imaged at its depth, Q-compensation
 * amplitude restoration and 1/Q scaling (tests/test_oracle2d.py).
 */
#include "../include/erm2d.h"

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kC6[7] = {0.0, 1.2213363647, -0.096931457519, 0.017447662353,
                       -0.0029672895159, 0.00035900539822, -0.00002184781161};
const double kC4[5] = {0.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0};

template <typename T>
std::vector<T> ricker(double freq, double dt_s) {  /* make_ricker_new, zzzzz:136-157 */
    const double PI = 3.141592653589793;
    T B = (T)(std::sqrt(6.0) / (PI * freq));
    int N = (int)std::ceil(1.35 * (double)B / dt_s);
    std::vector<T> w(2 * N + 1);
    w[N] = (T)1;
    for (int i = 1; i <= N; ++i) {
        T t = (T)dt_s * (T)i, u = (T)(2.0 * std::sqrt(6.0)) * t / B;
        w[N + i] = w[N - i] = (T)0.5 * ((T)2 - u * u) * std::exp(-u * u / (T)4);
    }
    return w;
}

struct Sls {  /* per cell: unrelaxed-modulus factors and memory coefficients */
    double fM, fU;             /* M_U = rho vp^2 fM, mu_U = rho vs^2 fU */
    double gP[3], gS[3];       /* kP_l = rho vp^2 gP_l, kS_l = rho vs^2 gS_l */
    double taus[3];
};

/* cuda_cal_viscoelastic (viscoelastic_equation.cu:43-56) for L = 1; tau-method for L = 3 */
Sls sls_params(int L, double freq, double qp, double qs) {
    Sls s{};
    const double w = 2 * 3.1415926 * freq;  /* pai as the reference */
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
    for (int l = 0; l < 3; ++l) {
        s.gP[l] = -tP / s.taus[l];
        s.gS[l] = -tS / s.taus[l];
    }
    return s;
}

template <typename T>
struct Oracle2D {
    erm2d_params p;
    int R = 4, L = 0, nnx = 0, nnz = 0;
    size_t N = 0;
    T c[7] = {0};
    T cx = 0, cz = 0, dtr = 0;
    T qsign = 1;  // -1 during a Q-compensated receiver pass
    enum { VX, VZ, XX, ZZ, XZ, TP, VXP, VZP, NF };
    std::vector<T> f[NF];
    std::vector<T> r[3][4];  /* memory: p, xx, zz, xz per mechanism */
    struct Model {
        std::vector<T> M, U, B, kP[3], kS[3], decay[3];  /* decay_l = dt/tau_sig_l */
        bool have = false;
    } mod[2];
    std::vector<T> att;
    std::vector<T> wavelet;
    long src = 0;
    std::vector<long> recs;
    std::vector<int> ex_time;
    std::vector<T> ex_amp, ex_tp, ex_vxp, ex_vzp, img_tp, img_pp, img_ps;
    std::vector<T> r_vxp, r_vzp;  // receiver vector P at each cell's imaging step
    std::vector<int> ex_tp_time;   // tp excitation (caculate_ex_tp_time_new)
    std::vector<T> ex_tp_amp;
    std::vector<T> vp_m[2], vs_m[2];  // loaded vp, vs (decom_new)
    /* C-PML (no reference counterpart): coefficients a, b per axis at integer
     * ("full") and staggered ("half": x - 1/2, z + 1/2) positions and 10
     * memory variables on the whole grid (zero outside the layers):
     * stress side 0 Dx vx (x full), 1 Dx vz (x half), 2 Dz vz (z full),
     * 3 Dz vx (z half); velocity side 4 Dx sxx (x half), 5 Dz sxz (z full),
     * 6 Dx sxz (x full), 7 Dz szz (z half), 8 Dx tp (x half), 9 Dz tp (z half). */
    bool cpml = false;
    double vmax_loaded = 0;
    std::vector<T> ca[2][2], cb[2][2];  /* [axis 0 x, 1 z][0 full, 1 half] */
    std::vector<T> psi[10];
    int last_model = 1;                // model of the last pass (0 true, 1 background)
    std::vector<T> rec;  /* [2][lt][nrec] */

    size_t id(int i, int k) const { return (size_t)i * nnz + k; }
    bool interior(int i, int k) const { return i >= R && i < nnx - R && k >= R && k < nnz - R; }

    int init(const erm2d_params* pp) {
        p = *pp;
        R = p.fd_order == 12 ? 6 : (p.fd_order == 8 ? 4 : -1);
        if (R < 0) return -1;
        L = p.nsls;
        if (L != 0 && L != 1 && L != 3) return -1;
        nnx = p.nx + p.bl + p.br;
        nnz = p.nz + p.bu + p.bd;
        if (nnx < 2 * R + 1 || nnz < 2 * R + 1) return -5;
        N = (size_t)nnx * nnz;
        for (int m = 0; m <= R; ++m) c[m] = (T)(R == 6 ? kC6[m] : kC4[m]);
        dtr = (T)(p.dt / 1000.0);
        cx = (T)(p.dt / 1000.0 / p.dx);
        cz = (T)(p.dt / 1000.0 / p.dz);
        for (auto& v : f) v.assign(N, 0);
        for (int l = 0; l < L; ++l) for (auto& v : r[l]) v.assign(N, 0);
        /* make_attenuation_new (zzzzz:238-325): coe * s^3, s = depth into the layer / (b-1) */
        std::vector<double> ax(nnx, 0), az(nnz, 0);
        auto face = [&](int b, bool rising, std::vector<double>& axis, int start) {
            for (int t = 0; t < b; ++t) {
                double s = rising ? t / (b - 1.0) : (b - 1.0 - t) / (b - 1.0);
                axis[start + t] += p.coe_attenuation * std::pow(s, 3.0);
            }
        };
        cpml = p.boundary == ERM_BOUNDARY_CPML;
        if (p.boundary != ERM_BOUNDARY_DAMP && !cpml) return -1;
        if (!cpml) {  /* C-PML replaces the damping */
            if (p.bu > 1) face(p.bu, false, az, 0);
            if (p.bd > 1) face(p.bd, true, az, p.bu + p.nz);
            if (p.bl > 1) face(p.bl, false, ax, 0);
            if (p.br > 1) face(p.br, true, ax, p.bl + p.nx);
        }
        att.assign(N, 0);
        for (int i = 0; i < nnx; ++i)
            for (int k = 0; k < nnz; ++k) att[id(i, k)] = (T)(ax[i] + az[k]);
        wavelet = ricker<T>(p.freq, p.dt / 1000.0);
        recs.clear();
        for (int q = 0; q < p.nrec; ++q) {
            int i = p.bl + p.rec_x0 + q * p.rec_dx, k = p.bu + p.rec_z;
            if (i < 0 || i >= nnx || k < 0 || k >= nnz) return -1;
            recs.push_back((long)id(i, k));
        }
        rec.assign((size_t)2 * p.lt * p.nrec, 0);
        ex_time.assign(N, -1);
        ex_tp_time.assign(N, -1);
        for (auto* v : {&ex_amp, &ex_tp, &ex_vxp, &ex_vzp, &img_tp, &img_pp, &img_ps, &r_vxp, &r_vzp, &ex_tp_amp})
            v->assign(N, 0);
        return set_shot(p.shot_x, p.shot_z);
    }

    int set_shot(int sx, int sz) {
        int i = p.bl + sx, k = p.bu + sz;
        if (!interior(i, k)) return -1;
        src = (long)id(i, k);
        return 0;
    }

    int load_model(int which, const float* vp, const float* vs, const float* rho, const float* qp, const float* qs) {
        Model& m = mod[which];
        m.M.assign(N, 0); m.U.assign(N, 0); m.B.assign(N, 0);
        vp_m[which].assign(N, 0); vs_m[which].assign(N, 0);
        for (int l = 0; l < L; ++l) { m.kP[l].assign(N, 0); m.kS[l].assign(N, 0); m.decay[l].assign(N, 0); }
        for (int i = 0; i < nnx; ++i)
            for (int k = 0; k < nnz; ++k) {
                /* edge-replicated padding (read_velocity_new, zzzzz:327-357) */
                int ix = std::min(std::max(i - p.bl, 0), p.nx - 1), iz = std::min(std::max(k - p.bu, 0), p.nz - 1);
                size_t s = (size_t)ix * p.nz + iz, n = id(i, k);
                double a = vp[s], b = vs[s], rr = rho[s];
                if (!(a > 0) || b < 0 || !(rr > 0)) return -1;
                vp_m[which][n] = (T)a; vs_m[which][n] = (T)b;
                Sls q = sls_params(L, p.freq, L ? qp[s] : 1.0, L ? qs[s] : 1.0);
                m.M[n] = (T)(rr * a * a * q.fM);
                m.U[n] = (T)(rr * b * b * q.fU);
                m.B[n] = (T)(1.0 / rr);
                for (int l = 0; l < L; ++l) {
                    m.kP[l][n] = (T)(rr * a * a * q.gP[l]);
                    m.kS[l][n] = (T)(rr * b * b * q.gS[l]);
                    m.decay[l][n] = (T)(p.dt / 1000.0 / q.taus[l]);
                }
            }
        m.have = true;
        for (size_t n = 0; n < N; ++n) vmax_loaded = std::max(vmax_loaded, (double)vp_m[which][n]);
        if (cpml) build_cpml();
        return 0;
    }

    /* C-PML as in 3D (Roden & Gedney 2000; Komatitsch & Martin 2007), kappa = 1:
     * psi' = b psi + a e, e' = e + psi'; b = exp(-(d+alpha) dt), a = d (b-1)/(d+alpha),
     * d = d0 r^2, d0 = -3 vmax ln(Rc) / (2 L h), alpha = pi f0 (1 - r), r = depth into
     * the layer / L, L = width - R (the outer R cells are never updated). */
    void build_cpml() {
        const double vmax = p.cpml_vmax > 0 ? p.cpml_vmax : vmax_loaded;
        const double Rc = p.cpml_r > 0 ? p.cpml_r : 1e-3;
        const double dts = p.dt / 1000.0;
        const int n2[2] = {nnx, nnz}, lo[2] = {p.bl, p.bu}, hi[2] = {p.br, p.bd};
        const double h[2] = {p.dx, p.dz}, half_shift[2] = {-0.5, +0.5};
        for (int a = 0; a < 2; ++a)
            for (int f = 0; f < 2; ++f) {
                ca[a][f].assign(n2[a], 0);
                cb[a][f].assign(n2[a], 1);
                for (int i = 0; i < n2[a]; ++i) {
                    double x = i + (f ? half_shift[a] : 0.0), s = 0, Lw = 0;
                    if (lo[a] > 0 && x < lo[a]) { s = lo[a] - x; Lw = std::max(lo[a] - R, 1); }
                    else if (hi[a] > 0 && x > n2[a] - hi[a] - 1) { s = x - (n2[a] - hi[a] - 1); Lw = std::max(hi[a] - R, 1); }
                    if (Lw <= 0) continue;
                    double r = s / Lw;
                    double d0 = -3.0 * vmax * std::log(Rc) / (2.0 * Lw * h[a]);
                    double d = d0 * r * r, al = M_PI * p.freq * (1.0 - std::min(r, 1.0));
                    double bb = std::exp(-(d + al) * dts);
                    cb[a][f][i] = (T)bb;
                    ca[a][f][i] = (T)(d > 0 ? d * (bb - 1.0) / (d + al) : 0.0);
                }
            }
        for (auto& v : psi) v.assign(N, (T)0);
    }
    inline T pml(int q, size_t n, int axis, int half, int idx, T e) {
        if (!cpml) return e;
        const T bq = cb[axis][half][idx], aq = ca[axis][half][idx];
        if (aq == (T)0) return e;
        T v = bq * psi[q][n] + aq * e;
        psi[q][n] = v;
        return e + v;
    }

    /* staggered operators on a z-fastest grid: D+ centred at +1/2, D- at -1/2 */
    inline T dp(const T* q, ptrdiff_t s) const { T a = 0; for (int m = 1; m <= R; ++m) a += c[m] * (q[m * s] - q[-(m - 1) * s]); return a; }
    inline T dm(const T* q, ptrdiff_t s) const { T a = 0; for (int m = 1; m <= R; ++m) a += c[m] * (q[(m - 1) * s] - q[-m * s]); return a; }
    T ab(size_t n, T* b) const { T h = att[n] * dtr / (T)2; *b = (T)1 - h; return (T)1 / ((T)1 + h); }

    /* strain increments E = (Vx, Vz, S) of a velocity field at cell n */
    void strain(const T* vx, const T* vz, size_t n, T* E) const {
        const ptrdiff_t sX = nnz;
        E[0] = cx * dp(vx + n, sX);
        E[1] = cz * dm(vz + n, 1);
        E[2] = cx * dm(vz + n, sX) + cz * dp(vx + n, 1);
    }
    /* stress forces F = (fx, fz) at cell n */
    void force(const T* xx, const T* zz, const T* xz, size_t n, T* F) const {
        const ptrdiff_t sX = nnz;
        F[0] = cx * dm(xx + n, sX) + cz * dm(xz + n, 1);
        F[1] = cx * dp(xz + n, sX) + cz * dp(zz + n, 1);
    }

    /* stress + memory update of state (fv, rv) with model m (viscoelastic_equation.cu:294-323) */
    void stress(std::vector<T>* fv, std::vector<T> (*rv)[4], const Model& m, T* Ehist = nullptr) {
#pragma omp parallel for schedule(static)
        for (int i = R; i < nnx - R; ++i)
            for (int k = R; k < nnz - R; ++k) {
                size_t n = id(i, k);
                T E[3];
                {  /* the same expression tree with or without C-PML (pml() passes e through when off) */
                    const ptrdiff_t sX = nnz;
                    T e0 = cx * dp(fv[VX].data() + n, sX), e1 = cz * dm(fv[VZ].data() + n, 1);
                    T e2x = cx * dm(fv[VZ].data() + n, sX), e2z = cz * dp(fv[VX].data() + n, 1);
                    E[0] = pml(0, n, 0, 0, i, e0);
                    E[1] = pml(2, n, 1, 0, k, e1);
                    E[2] = pml(1, n, 0, 1, i, e2x) + pml(3, n, 1, 1, k, e2z);
                }
                if (Ehist && i >= p.bl && i < p.bl + p.nx && k >= p.bu && k < p.bu + p.nz) {
                    size_t q = (size_t)(i - p.bl) * p.nz + (k - p.bu), mn = (size_t)p.nx * p.nz;
                    Ehist[q] = E[0]; Ehist[mn + q] = E[1]; Ehist[2 * mn + q] = E[2];
                }
                T b, a = ab(n, &b);
                T M = m.M[n], U = m.U[n], La = M - 2 * U;
                T sp = 0, sxx = 0, szz = 0, sxz = 0;
                for (int l = 0; l < L; ++l) {
                    T bl = b - m.decay[l][n], kP = m.kP[l][n], kS = m.kS[l][n];
                    T rp = a * (bl * rv[l][0][n] + kP * (E[0] + E[1]));
                    T rxx = a * (bl * rv[l][1][n] - 2 * kS * E[1]);
                    T rzz = a * (bl * rv[l][2][n] - 2 * kS * E[0]);
                    T rxz = a * (bl * rv[l][3][n] + kS * E[2]);
                    rv[l][0][n] = rp; rv[l][1][n] = rxx; rv[l][2][n] = rzz; rv[l][3][n] = rxz;
                    sp += rp; sxx += rxx; szz += rzz; sxz += rxz;
                }
                // qsign = -1: Q-compensated receiver propagation, the memory term's sign
                // flipped as in rfwd_txxzzxzpp_viscoelastic_and_memory (viscoelastic_equation.cu:1541-1569)
                const T dq = dtr * qsign;
                fv[XX][n] = a * (b * fv[XX][n] + M * E[0] + La * E[1] + dq * (sp + sxx));
                fv[ZZ][n] = a * (b * fv[ZZ][n] + La * E[0] + M * E[1] + dq * (sp + szz));
                fv[TP][n] = a * (b * fv[TP][n] + M * (E[0] + E[1]) + dq * sp);
                fv[XZ][n] = a * (b * fv[XZ][n] + U * E[2] + dq * sxz);
            }
    }

    void velocity(std::vector<T>* fv, const Model& m, int mode, int it, T* Fhist = nullptr) {
        const ptrdiff_t sX = nnz;
#pragma omp parallel for schedule(static)
        for (int i = R; i < nnx - R; ++i)
            for (int k = R; k < nnz - R; ++k) {
                size_t n = id(i, k);
                T F[2];
                T px = cx * dm(fv[TP].data() + n, sX), pz = cz * dp(fv[TP].data() + n, 1);
                {
                    T f0x = cx * dm(fv[XX].data() + n, sX), f0z = cz * dm(fv[XZ].data() + n, 1);
                    T f1x = cx * dp(fv[XZ].data() + n, sX), f1z = cz * dp(fv[ZZ].data() + n, 1);
                    F[0] = pml(4, n, 0, 1, i, f0x) + pml(5, n, 1, 0, k, f0z);
                    F[1] = pml(6, n, 0, 0, i, f1x) + pml(7, n, 1, 1, k, f1z);
                    px = pml(8, n, 0, 1, i, px);
                    pz = pml(9, n, 1, 1, k, pz);
                }
                if (Fhist && i >= p.bl && i < p.bl + p.nx && k >= p.bu && k < p.bu + p.nz) {
                    size_t q = (size_t)(i - p.bl) * p.nz + (k - p.bu), mn = (size_t)p.nx * p.nz;
                    Fhist[q] = F[0]; Fhist[mn + q] = F[1];
                }
                T b, a = ab(n, &b), bo = m.B[n];
                T nvx = a * (b * fv[VX][n] + bo * F[0]);
                T nvz = a * (b * fv[VZ][n] + bo * F[1]);
                T nxp = a * (b * fv[VXP][n] + bo * px);
                T nzp = a * (b * fv[VZP][n] + bo * pz);
                fv[VX][n] = nvx; fv[VZ][n] = nvz; fv[VXP][n] = nxp; fv[VZP][n] = nzp;
            }
        if (mode == 1) excite(fv, it);
        else if (mode == 2) image(fv, it);
    }

    /* vzp at cell n: own value, or the mean over (i, k+-1), (i+-1, k) of the
     * _new variants (caculate_ex_x_z_new, elastic_2D_kernel_3.cu:91-109) */
    T vzp_at(const T* a, size_t n) const {
        if (!p.ex_vzp_avg) return a[n];
        return (a[n + 1] + a[n - 1] + a[n + nnz] + a[n - nnz]) / (T)4;
    }
    /* excitation after the velocity update of step it (interior cells):
     * caculate_ex_amp_time_new (elastic_2D_kernel_3.cu:30-48): ex_time = it where
     * |vP| = sqrt(vxp^2 + vzp^2) exceeds ex_amp (strict '<': first maximum), with
     * tp and the vP components (caculate_ex_x_z[_new], :71-109) of that step;
     * caculate_ex_tp_time_new (:50-69): ex_tp_time = it - ex_tp_shift where |tp|
     * exceeds |ex_tp_amp|, ex_tp_amp = tp (signed). */
    void excite(std::vector<T>* fv, int it) {
        std::vector<T> zq(N);
        for (int i = R; i < nnx - R; ++i)
            for (int k = R; k < nnz - R; ++k) zq[id(i, k)] = vzp_at(fv[VZP].data(), id(i, k));
        for (int i = R; i < nnx - R; ++i)
            for (int k = R; k < nnz - R; ++k) {
                size_t n = id(i, k);
                T nxp = fv[VXP][n], nzp = fv[VZP][n];
                T amp = std::sqrt(nxp * nxp + nzp * nzp);
                if (ex_amp[n] < amp) {
                    ex_time[n] = it; ex_amp[n] = amp; ex_tp[n] = fv[TP][n];
                    ex_vxp[n] = nxp; ex_vzp[n] = zq[n];
                }
                T tpv = fv[TP][n];
                if (std::fabs(ex_tp_amp[n]) < std::fabs(tpv)) { ex_tp_time[n] = it - p.ex_tp_shift; ex_tp_amp[n] = tpv; }
            }
    }
    /* imaging of step it (interior cells; imaging_correlation_ex_2D and
     * imaging_inner_product_ex_2D[_new], elastic_2D_kernel_3.cu:542-605):
     * where ex_tp_time == it, I_tp = tp / ex_tp_amp; where ex_time == it,
     * I_PP = (ex_vxp vxp + ex_vzp vzp') / |ex|^2 and I_PS likewise with
     * vS = v - vP, vzp' and vzs' averaged as in the excitation; a zero
     * denominator images 0.  Also keeps the receiver vP of the imaging step. */
    void image(std::vector<T>* fv, int it) {
        for (int i = R; i < nnx - R; ++i)
            for (int k = R; k < nnz - R; ++k) {
                size_t n = id(i, k);
                if (ex_tp_time[n] == it) {
                    T a = ex_tp_amp[n];
                    img_tp[n] = a != (T)0 ? fv[TP][n] / a : (T)0;
                }
                if (ex_time[n] != it) continue;
                T nxp = fv[VXP][n], nzp = vzp_at(fv[VZP].data(), n);
                T nxs = fv[VX][n] - fv[VXP][n], nzs = vzp_at(fv[VZ].data(), n) - nzp;
                T ex = ex_vxp[n], ez = ex_vzp[n], den = ex * ex + ez * ez;
                img_pp[n] = den > 0 ? (ex * nxp + ez * nzp) / den : (T)0;
                img_ps[n] = den > 0 ? (ex * nxs + ez * nzs) / den : (T)0;
                r_vxp[n] = fv[VXP][n]; r_vzp[n] = fv[VZP][n];
            }
    }

    void add_source(std::vector<T>* fv, int it) {
        if (it >= 0 && it < (int)wavelet.size()) { fv[XX][src] += wavelet[it]; fv[ZZ][src] += wavelet[it]; }
    }

    int step(int pass, int it_begin, int nsteps) {
        const Model& m = mod[pass == 0 ? 0 : 1];
        if (!m.have) return -4;
        last_model = pass == 0 ? 0 : 1;
        for (int s = 0; s < nsteps; ++s) {
            if (pass == 2) {
                int it = it_begin - s;
                if (it < 0 || it >= p.lt) return -1;
                for (int q = 0; q < p.nrec; ++q) {
                    f[VX][recs[q]] += rec[((size_t)0 * p.lt + it) * p.nrec + q];
                    f[VZ][recs[q]] += rec[((size_t)1 * p.lt + it) * p.nrec + q];
                }
                qsign = p.qcomp ? (T)-1 : (T)1;
                stress(f, r, m);
                qsign = 1;
                velocity(f, m, 2, it);
            } else {
                int it = it_begin + s;
                add_source(f, it);
                stress(f, r, m);
                velocity(f, m, pass, it);
                if (pass == 0 && it < p.lt)
                    for (int q = 0; q < p.nrec; ++q) {
                        rec[((size_t)0 * p.lt + it) * p.nrec + q] = f[VX][recs[q]];
                        rec[((size_t)1 * p.lt + it) * p.nrec + q] = f[VZ][recs[q]];
                    }
            }
        }
        return 0;
    }

    void zero() {
        for (auto& v : f) std::fill(v.begin(), v.end(), (T)0);
        for (auto& v : psi) std::fill(v.begin(), v.end(), (T)0);
        for (int l = 0; l < L; ++l) for (auto& v : r[l]) std::fill(v.begin(), v.end(), (T)0);
    }

    /* decom (elastic_2D_kernel_2.cu:621-666): divergence at the stress node,
     * curl at the sxz node; interior cells. */
    void decom(std::vector<T>& dv, std::vector<T>& cu) const {
        dv.assign(N, 0); cu.assign(N, 0);
        const ptrdiff_t sX = nnz;
        for (int i = R; i < nnx - R; ++i)
            for (int k = R; k < nnz - R; ++k) {
                size_t n = id(i, k);
                dv[n] = (T)(1.0 / p.dx) * dp(f[VX].data() + n, sX) + (T)(1.0 / p.dz) * dm(f[VZ].data() + n, 1);
                cu[n] = (T)(1.0 / p.dz) * dp(f[VX].data() + n, 1) - (T)(1.0 / p.dx) * dm(f[VZ].data() + n, sX);
            }
    }

    /* decom_new (elastic_2D_kernel_2.cu:671-791): the decom stencils scaled by
     * vp at the stress node and by the mean of vs over (i,k), (i,k+1), (i+1,k),
     * (i+1,k+1) (:726, :758, :789); model `which`. */
    void decom_new(int which, std::vector<T>& dv, std::vector<T>& cu) const {
        decom(dv, cu);
        const std::vector<T>& a = vp_m[which];
        const std::vector<T>& b = vs_m[which];
        for (int i = R; i < nnx - R; ++i)
            for (int k = R; k < nnz - R; ++k) {
                size_t n = id(i, k);
                dv[n] *= a[n];
                cu[n] *= (b[n] + b[n + 1] + b[n + nnz] + b[n + 1 + nnz]) / (T)4;
            }
    }

    /* ------------------------------------------------------------ LSRTM */
    struct Pert { std::vector<T> mp, ms, mr; };
    Pert unpack(const float* dm) const {
        Pert P;
        size_t mn = (size_t)p.nx * p.nz;
        P.mp.assign(dm, dm + mn); P.ms.assign(dm + mn, dm + 2 * mn); P.mr.assign(dm + 2 * mn, dm + 3 * mn);
        return P;
    }
    bool model_cell(int i, int k, size_t* q) const {
        if (i < p.bl || i >= p.bl + p.nx || k < p.bu || k >= p.bu + p.nz) return false;
        *q = (size_t)(i - p.bl) * p.nz + (k - p.bu);
        return true;
    }

    /* Relative perturbations of the P and S moduli (rho vp^2, rho vs^2) from
     * the three inverted parameters (inversion_para, the mapping of
     * cuda_cal_dem_parameter_viscoelastic_media_new, viscoelastic_equation.cu:
     * 2431-2486, written for relative perturbations): 1 (lambda, mu, rho):
     * sP = (lambda m1 + 2 mu m2) / (lambda + 2 mu), sS = m2; 2 (vp, vs, rho):
     * sP = 2 m1 + m3, sS = 2 m2 + m3; 3 (Ip, Is, rho): sP = 2 m1 - m3,
     * sS = 2 m2 - m3.  The density perturbation is m3 in all three. */
    int para() const { return p.inversion_para == 1 || p.inversion_para == 3 ? p.inversion_para : 2; }
    void rel_moduli(size_t n, T m1, T m2, T m3, T* sP, T* sS) const {
        if (para() == 1) {
            const T a = vp_m[1][n], b = vs_m[1][n];
            *sP = ((a * a - 2 * b * b) * m1 + 2 * b * b * m2) / (a * a);
            *sS = m2;
        } else if (para() == 3) {
            *sP = 2 * m1 - m3;
            *sS = 2 * m2 - m3;
        } else {
            *sP = 2 * m1 + m3;
            *sS = 2 * m2 + m3;
        }
    }

    /* d = L dm: background and Born wavefields stepped together. */
    int born(const float* dmv, float* d) {
        if (cpml) return -1;  /* the LSRTM operators use the damping boundary */
        const Model& m = mod[1];
        if (!m.have) return -4;
        Pert P = unpack(dmv);
        size_t mn = (size_t)p.nx * p.nz;
        std::vector<T> Eh(3 * mn), Fh(2 * mn);
        std::vector<T> g[NF];
        std::vector<T> gr[3][4];
        for (auto& v : g) v.assign(N, 0);
        for (int l = 0; l < L; ++l) for (auto& v : gr[l]) v.assign(N, 0);
        zero();
        for (int it = 0; it < p.lt; ++it) {
            add_source(f, it);
            stress(f, r, m, Eh.data());
            stress(g, gr, m);
            /* Born stress/memory sources: cuda_cal_dem_parameter_viscoelastic_media_new (:2398-2505) */
#pragma omp parallel for schedule(static)
            for (int i = R; i < nnx - R; ++i)
                for (int k = R; k < nnz - R; ++k) {
                    size_t q;
                    if (!model_cell(i, k, &q)) continue;
                    size_t n = id(i, k);
                    T b, a = ab(n, &b);
                    T E0 = Eh[q], E1 = Eh[mn + q], E2 = Eh[2 * mn + q];
                    T sP, sS;
                    rel_moduli(n, P.mp[q], P.ms[q], P.mr[q], &sP, &sS);
                    T dM = m.M[n] * sP, dU = m.U[n] * sS, dL = dM - 2 * dU;
                    T sp = 0, sxx = 0, szz = 0, sxz = 0;
                    for (int l = 0; l < L; ++l) {
                        T dkP = m.kP[l][n] * sP, dkS = m.kS[l][n] * sS;
                        T rp = a * dkP * (E0 + E1), rxx = a * (-2 * dkS * E1), rzz = a * (-2 * dkS * E0), rxz = a * dkS * E2;
                        gr[l][0][n] += rp; gr[l][1][n] += rxx; gr[l][2][n] += rzz; gr[l][3][n] += rxz;
                        sp += rp; sxx += rxx; szz += rzz; sxz += rxz;
                    }
                    g[XX][n] += a * (dM * E0 + dL * E1 + dtr * (sp + sxx));
                    g[ZZ][n] += a * (dL * E0 + dM * E1 + dtr * (sp + szz));
                    g[XZ][n] += a * (dU * E2 + dtr * sxz);
                }
            velocity(f, m, 0, it, Fh.data());
            velocity(g, m, 0, it);
#pragma omp parallel for schedule(static)
            for (int i = R; i < nnx - R; ++i)
                for (int k = R; k < nnz - R; ++k) {
                    size_t q;
                    if (!model_cell(i, k, &q)) continue;
                    size_t n = id(i, k);
                    T b, a = ab(n, &b);
                    T s = -a * m.B[n] * P.mr[q];  /* delta(1/rho) = -mr/rho: -drho dv/dt (:2488-2489) */
                    g[VX][n] += s * Fh[q];
                    g[VZ][n] += s * Fh[mn + q];
                }
            for (int qq = 0; qq < p.nrec; ++qq) {
                d[((size_t)0 * p.lt + it) * p.nrec + qq] = (float)g[VX][recs[qq]];
                d[((size_t)1 * p.lt + it) * p.nrec + qq] = (float)g[VZ][recs[qq]];
            }
        }
        return 0;
    }

    /* g = L^T r: exact transpose of born(), stepping back from lt-1.  EHin /
     * FHin: the background's E and F histories computed elsewhere (the strips
     * reconstruction, adjoint_strips); else the background pass stores them. */
    int adjoint(const float* rr, float* gout, const std::vector<T>* EHin = nullptr,
                const std::vector<T>* FHin = nullptr) {
        if (cpml) return -1;
        const Model& m = mod[1];
        if (!m.have) return -4;
        size_t mn = (size_t)p.nx * p.nz;
        /* background history: E(v0^n) and F(s0^{n+1}) over the model region */
        std::vector<T> EHl, FHl;
        if (!EHin || !FHin) {
            EHl.assign((size_t)p.lt * 3 * mn, 0);
            FHl.assign((size_t)p.lt * 2 * mn, 0);
            zero();
            for (int it = 0; it < p.lt; ++it) {
                add_source(f, it);
                stress(f, r, m, EHl.data() + (size_t)it * 3 * mn);
                velocity(f, m, 0, it, FHl.data() + (size_t)it * 2 * mn);
            }
        }
        const std::vector<T>& EH = EHin && FHin ? *EHin : EHl;
        const std::vector<T>& FH = EHin && FHin ? *FHin : FHl;
        std::vector<T> lv[2], ls[3], lr[3][4], W[3];
        for (auto& v : lv) v.assign(N, 0);
        for (auto& v : ls) v.assign(N, 0);
        for (auto& v : W) v.assign(N, 0);
        for (int l = 0; l < L; ++l) for (auto& v : lr[l]) v.assign(N, 0);
        std::vector<T> gp(mn, 0), gs(mn, 0), gr_(mn, 0);
        std::vector<T> Z[2];
        for (auto& v : Z) v.assign(N, 0);
        const ptrdiff_t sX = nnz;
        for (int it = p.lt - 1; it >= 0; --it) {
            const T* Eh = EH.data() + (size_t)it * 3 * mn;
            const T* Fh = FH.data() + (size_t)it * 2 * mn;
            for (int q = 0; q < p.nrec; ++q) {  /* record transpose */
                lv[0][recs[q]] += (T)rr[((size_t)0 * p.lt + it) * p.nrec + q];
                lv[1][recs[q]] += (T)rr[((size_t)1 * p.lt + it) * p.nrec + q];
            }
            /* V^T: ls += F^T(a B lv); lv <- a b lv; density gradient */
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nnx; ++i)
                for (int k = 0; k < nnz; ++k) {
                    size_t n = id(i, k);
                    if (!interior(i, k)) { Z[0][n] = Z[1][n] = 0; continue; }
                    T b, a = ab(n, &b);
                    Z[0][n] = a * m.B[n] * lv[0][n];
                    Z[1][n] = a * m.B[n] * lv[1][n];
                    size_t q;
                    if (model_cell(i, k, &q)) gr_[q] += -(Fh[q] * Z[0][n] + Fh[mn + q] * Z[1][n]);
                }
#pragma omp parallel for schedule(static)
            for (int i = R; i < nnx - R; ++i)
                for (int k = R; k < nnz - R; ++k) {
                    size_t n = id(i, k);
                    ls[0][n] += -cx * dp(Z[0].data() + n, sX);                          /* (D-x)^T = -D+x */
                    ls[2][n] += -cz * dp(Z[0].data() + n, 1) - cx * dm(Z[1].data() + n, sX);
                    ls[1][n] += -cz * dm(Z[1].data() + n, 1);                           /* (D+z)^T = -D-z */
                }
#pragma omp parallel for schedule(static)
            for (int i = R; i < nnx - R; ++i)
                for (int k = R; k < nnz - R; ++k) {
                    size_t n = id(i, k);
                    T b, a = ab(n, &b);
                    lv[0][n] *= a * b;
                    lv[1][n] *= a * b;
                }
            /* S^T: mu_s = a ls, mu_r = lr + dt P^T mu_s; W = C mu_s + a K^T mu_r; gradients */
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nnx; ++i)
                for (int k = 0; k < nnz; ++k) {
                    size_t n = id(i, k);
                    if (!interior(i, k)) { W[0][n] = W[1][n] = W[2][n] = 0; continue; }
                    T b, a = ab(n, &b);
                    T mxx = a * ls[0][n], mzz = a * ls[1][n], mxz = a * ls[2][n];
                    T M = m.M[n], U = m.U[n], La = M - 2 * U;
                    T wx = M * mxx + La * mzz, wz = La * mxx + M * mzz, ws = U * mxz;
                    size_t q;
                    bool inm = model_cell(i, k, &q);
                    T E0 = 0, E1 = 0, E2 = 0;
                    if (inm) { E0 = Eh[q]; E1 = Eh[mn + q]; E2 = Eh[2 * mn + q]; }
                    T gkp = 0, gks = 0;
                    for (int l = 0; l < L; ++l) {
                        T rp = lr[l][0][n] + dtr * (mxx + mzz), rxx = lr[l][1][n] + dtr * mxx;
                        T rzz = lr[l][2][n] + dtr * mzz, rxz = lr[l][3][n] + dtr * mxz;
                        T kP = m.kP[l][n], kS = m.kS[l][n], bl = b - m.decay[l][n];
                        wx += a * (kP * rp - 2 * kS * rzz);
                        wz += a * (kP * rp - 2 * kS * rxx);
                        ws += a * kS * rxz;
                        if (inm) {
                            gkp += kP * (E0 + E1) * a * rp;
                            gks += kS * (-2 * E1 * a * rxx - 2 * E0 * a * rzz + E2 * a * rxz);
                        }
                        lr[l][0][n] = a * bl * rp; lr[l][1][n] = a * bl * rxx;
                        lr[l][2][n] = a * bl * rzz; lr[l][3][n] = a * bl * rxz;
                    }
                    W[0][n] = wx; W[1][n] = wz; W[2][n] = ws;
                    if (inm) {
                        T gM = (E0 + E1) * (mxx + mzz), gU = -2 * (E1 * mxx + E0 * mzz) + E2 * mxz;
                        gp[q] += 2 * M * gM + 2 * gkp;
                        gs[q] += 2 * U * gU + 2 * gks;
                        gr_[q] += M * gM + U * gU + gkp + gks;
                    }
                    ls[0][n] = b * mxx; ls[1][n] = b * mzz; ls[2][n] = b * mxz;
                }
#pragma omp parallel for schedule(static)
            for (int i = R; i < nnx - R; ++i)
                for (int k = R; k < nnz - R; ++k) {
                    size_t n = id(i, k);
                    lv[0][n] += -cx * dm(W[0].data() + n, sX) - cz * dm(W[2].data() + n, 1);
                    lv[1][n] += -cz * dp(W[1].data() + n, 1) - cx * dp(W[2].data() + n, sX);
                }
        }
        /* gp = 2 GP, gs = 2 GS, gr_ = GP + GS + GD in the (vp, vs, rho) parameters, GP / GS the
         * gradients of the relative P / S modulus perturbation and GD the direct density part;
         * the chain rule of rel_moduli gives the other parameterisations */
        for (size_t q = 0; q < mn; ++q) {
            const int i = (int)(q / p.nz) + p.bl, k = (int)(q % p.nz) + p.bu;
            const T GP = gp[q] / 2, GS = gs[q] / 2, GD = gr_[q] - GP - GS;
            T g1 = gp[q], g2 = gs[q], g3 = gr_[q];
            if (para() == 1) {
                const size_t n = id(i, k);
                const T a = vp_m[1][n], b = vs_m[1][n];
                g1 = GP * (a * a - 2 * b * b) / (a * a);
                g2 = GP * 2 * b * b / (a * a) + GS;
                g3 = GD;
            } else if (para() == 3) {
                g3 = GD - GP - GS;
            }
            gout[q] = (float)g1; gout[mn + q] = (float)g2; gout[2 * mn + q] = (float)g3;
        }
        return 0;
    }
};

/* ---------------------------------------------------------------------------
 * Reference-mode adjoint: the reference's continuous-adjoint state equations
 * and gradient, restated (adjoint_mode = 1; the exact transpose above is the
 * default).  Adjoint state: velocities av (x, z), strain-like stresses as
 * (xx, zz, xz) and per mechanism memory variables ar (xx, zz, xz).  One
 * reversed step, every update in the damping bracket a[b u + ...]:
 *   as += E(av)                 the strain increments of av (the kernel that
 *                               updates the adjoint stresses is absent from the
 *                               reference: no moduli, the positions of
 *                               viscoelastic_equation.cu:137-176)
 *   ar_l = a[(b - dt/tau_l) ar_l + dt as]   adjoint_fwd_memory_new (:2190-2196)
 *   avx  = a[b avx + Mp cx D-x as_xx + (Mp - 2Ms) cx D-x as_zz + Ms cz D-z as_xz
 *          + sum_l P_l cx D-x ar_xx + (P_l - 2S_l) cx D-x ar_zz + S_l cz D-z ar_xz]
 *                               adjoint_fwd_vx_viscoelastic (:1923-1971)
 *   avz  = a[b avz + Mp cz D+z as_zz + (Mp - 2Ms) cz D+z as_xx + Ms cx D+x as_xz
 *          + sum_l (P_l - 2S_l) cz D+z ar_xx + P_l cz D+z ar_zz + S_l cx D+x ar_xz]
 *                               adjoint_fwd_vz_viscoelastic (:2071-2120)
 * with the moduli of the cell (no averaging) Mp = vp^2 tau_eps/tau_sig,
 * P_l = vp^2 (1 - tau_eps/tau_sig)/tau_sig (:1908-1914), i.e. M/rho, U/rho,
 * kP_l/rho, kS_l/rho of the forward model.  The residual records are added
 * to av at the receivers before each reversed step (it = lt-1 .. 0).  The
 * gradient of cal_gradient_in_viscoelastic_media_new (:2248-2295) is
 * accumulated after the step over the model region from the background's
 * strain rates vx_x = E0/dt, vz_z = E1/dt, vx_z + vz_x = E2/dt and
 * accelerations v_t = B F/dt of step it:
 *   g_rho    += avx vx_t + avz vz_t
 *   g_lambda -= [mp (as_xx + as_zz) + sum_l np_l (ar_xx + ar_zz)] (vx_x + vz_z)
 *   g_mu     -= 2A + 2B + C + sum_l (2D_l + 2E_l + F_l)
 * A = (mp vx_x + (mp - ms) vz_z) as_xx, B = (mp vz_z + (mp - ms) vx_x) as_zz,
 * C = ms (vx_z + vz_x) as_xz, D, E, F the same with np_l, ns_l and ar_l;
 * mp = tau_eps/tau_sig = M/(rho vp^2), np_l = (1 - mp)/tau_sig = kP_l/(rho vp^2)
 * (ms, ns_l likewise with vs; 0 where vs = 0).  The adjoint tp and r_p of the
 * reference feed neither the velocities nor the gradient and are not kept.
 * ------------------------------------------------------------------------- */
template <typename T>
struct RefAdjoint {
    Oracle2D<T>& o;
    std::vector<T> av[2], as[3], ar[3][3];
    explicit RefAdjoint(Oracle2D<T>& oo) : o(oo) {
        for (auto& v : av) v.assign(o.N, 0);
        for (auto& v : as) v.assign(o.N, 0);
        for (int l = 0; l < o.L; ++l) for (auto& v : ar[l]) v.assign(o.N, 0);
    }
    void step(const typename Oracle2D<T>::Model& m) {
        const int R = o.R, nnx = o.nnx, nnz = o.nnz;
        const ptrdiff_t sX = nnz;
#pragma omp parallel for schedule(static)
        for (int i = R; i < nnx - R; ++i)
            for (int k = R; k < nnz - R; ++k) {
                const size_t n = o.id(i, k);
                T E[3];
                o.strain(av[0].data(), av[1].data(), n, E);
                T b, a = o.ab(n, &b);
                as[0][n] = a * (b * as[0][n] + E[0]);
                as[1][n] = a * (b * as[1][n] + E[1]);
                as[2][n] = a * (b * as[2][n] + E[2]);
                for (int l = 0; l < o.L; ++l) {
                    const T bl = b - m.decay[l][n];
                    for (int q = 0; q < 3; ++q) ar[l][q][n] = a * (bl * ar[l][q][n] + o.dtr * as[q][n]);
                }
            }
#pragma omp parallel for schedule(static)
        for (int i = R; i < nnx - R; ++i)
            for (int k = R; k < nnz - R; ++k) {
                const size_t n = o.id(i, k);
                T b, a = o.ab(n, &b);
                const T Bn = m.B[n], Mp = m.M[n] * Bn, Ms = m.U[n] * Bn;
                T fx = Mp * o.cx * o.dm(as[0].data() + n, sX) + (Mp - 2 * Ms) * o.cx * o.dm(as[1].data() + n, sX) +
                       Ms * o.cz * o.dm(as[2].data() + n, 1);
                T fz = Mp * o.cz * o.dp(as[1].data() + n, 1) + (Mp - 2 * Ms) * o.cz * o.dp(as[0].data() + n, 1) +
                       Ms * o.cx * o.dp(as[2].data() + n, sX);
                for (int l = 0; l < o.L; ++l) {
                    const T P = m.kP[l][n] * Bn, S = m.kS[l][n] * Bn;
                    fx += P * o.cx * o.dm(ar[l][0].data() + n, sX) + (P - 2 * S) * o.cx * o.dm(ar[l][1].data() + n, sX) +
                          S * o.cz * o.dm(ar[l][2].data() + n, 1);
                    fz += (P - 2 * S) * o.cz * o.dp(ar[l][0].data() + n, 1) + P * o.cz * o.dp(ar[l][1].data() + n, 1) +
                          S * o.cx * o.dp(ar[l][2].data() + n, sX);
                }
                av[0][n] = a * (b * av[0][n] + fx);
                av[1][n] = a * (b * av[1][n] + fz);
            }
    }
    /* gradient of one step (cal_gradient_in_viscoelastic_media_new), E/F of the step */
    void gradient(const typename Oracle2D<T>::Model& m, const T* Eh, const T* Fh, T* gl, T* gm, T* gr) {
        const size_t mn = (size_t)o.p.nx * o.p.nz;
        const T idt = (T)1 / o.dtr;
#pragma omp parallel for schedule(static)
        for (long q = 0; q < (long)mn; ++q) {
            const int i = (int)(q / o.p.nz) + o.p.bl, k = (int)(q % o.p.nz) + o.p.bu;
            if (!o.interior(i, k)) continue;
            const size_t n = o.id(i, k);
            const T vxx = Eh[q] * idt, vzz = Eh[mn + q] * idt, vxz = Eh[2 * mn + q] * idt;
            const T vxt = m.B[n] * Fh[q] * idt, vzt = m.B[n] * Fh[mn + q] * idt;
            const T rho = (T)1 / m.B[n], vp = o.vp_m[1][n], vs = o.vs_m[1][n];
            const T mp = m.M[n] / (rho * vp * vp), ms = vs > 0 ? m.U[n] / (rho * vs * vs) : (T)0;
            gr[q] += av[0][n] * vxt + av[1][n] * vzt;
            T lam = mp * (as[0][n] + as[1][n]);
            T A = (mp * vxx + (mp - ms) * vzz) * as[0][n], B = (mp * vzz + (mp - ms) * vxx) * as[1][n];
            T C = ms * vxz * as[2][n], DEF = 0;
            for (int l = 0; l < o.L; ++l) {
                const T np = m.kP[l][n] / (rho * vp * vp), ns = vs > 0 ? m.kS[l][n] / (rho * vs * vs) : (T)0;
                lam += np * (ar[l][0][n] + ar[l][1][n]);
                DEF += 2 * (np * vxx + (np - ns) * vzz) * ar[l][0][n] + 2 * (np * vzz + (np - ns) * vxx) * ar[l][1][n] +
                       ns * vxz * ar[l][2][n];
            }
            gl[q] -= lam * (vxx + vzz);
            gm[q] -= 2 * A + 2 * B + C + DEF;
        }
    }
};

/* reference-mode adjoint operator: background pass storing E/F, then the
 * reversed adjoint steps from the residual records; g = (g_lambda, g_mu, g_rho)
 * over the model region, the raw sums of the reference's gradient kernel */
template <typename T>
int adjoint_ref(Oracle2D<T>& o, const float* rr, float* gout) {
    if (o.cpml) return -1;
    const auto& m = o.mod[1];
    if (!m.have) return -4;
    const auto& p = o.p;
    const size_t mn = (size_t)p.nx * p.nz;
    std::vector<T> EH((size_t)p.lt * 3 * mn), FH((size_t)p.lt * 2 * mn);
    o.zero();
    for (int it = 0; it < p.lt; ++it) {
        o.add_source(o.f, it);
        o.stress(o.f, o.r, m, EH.data() + (size_t)it * 3 * mn);
        o.velocity(o.f, m, 0, it, FH.data() + (size_t)it * 2 * mn);
    }
    RefAdjoint<T> A(o);
    std::vector<T> gl(mn, 0), gm(mn, 0), gr(mn, 0);
    for (int it = p.lt - 1; it >= 0; --it) {
        for (int q = 0; q < p.nrec; ++q) {
            A.av[0][o.recs[q]] += (T)rr[((size_t)0 * p.lt + it) * p.nrec + q];
            A.av[1][o.recs[q]] += (T)rr[((size_t)1 * p.lt + it) * p.nrec + q];
        }
        A.step(m);
        A.gradient(m, EH.data() + (size_t)it * 3 * mn, FH.data() + (size_t)it * 2 * mn, gl.data(), gm.data(), gr.data());
    }
    for (size_t q = 0; q < mn; ++q) { gout[q] = (float)gl[q]; gout[mn + q] = (float)gm[q]; gout[2 * mn + q] = (float)gr[q]; }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Cross-correlation RTM with a choice of source-wavefield strategy and the
 * reference's PS polarity correction.
 *
 * Reversed step (rfwd_txxzzxzpp_viscoelastic_and_memory_old,
 * viscoelastic_equation.cu:1298-1435: the derivative and memory terms of the
 * stress update negated, the memory variables updated as in the forward
 * step; the velocity counterpart, absent from the reference, negates the
 * stress-divergence term): undoes forward step it from (v^{it+1}, s^{it}):
 *   v  = a[b v  - B F(s)],  vP = a[b vP - B grad tp]            (velocity first)
 *   r' = a[(b - dt/tau) r + K E(v)],  s = a[b (s - w_it) - C E(v) - dt r'] with
 *   the wavelet of step it taken off sxx, szz where the forward step adds it.
 * In the undamped model region of an elastic medium this is the exact inverse.
 *
 * Strategies (erm2d_rtm_xcorr):
 *   0 history  - the source pass stores vxp, vzp and div v of the model region
 *                after every step (save_all_wavefield / save_and_set_wavefiled,
 *                elastic_2D_kernel_2.cu:793-827, viscoelastic_equation.cu:3-21);
 *   1 reverse  - the source state at the end of the source pass is stepped
 *                backward with the reversed step alone (the absorbing boundary
 *                keeps absorbing, so the reconstruction is approximate);
 *   2 strips   - as 1, but the fields of a ring of width R around the model
 *                region shrunk by one cell are stored every source step and
 *                restored before each reversed half step: exact for elastic
 *                media (no reference counterpart; nsls = 0 only).
 * Receiver pass at it = lt-1 .. 0: inject, stress, velocity, then image with
 * the source wavefield after forward step it (history it, or the reversed
 * state before undoing step it), for it < lt_forward.
 * Images over the model region, accumulated over the steps:
 *   0 PP  = sum vP_s . vP_r              1 PS  = sum vP_s . (v_r - vP_r)
 *   2 PPd = sum div_s div_r              3 PSc = sum div_s curl_r
 * The source fields enter only over the model region (zero outside it, as
 * the history of save_and_set_wavefiled holds them), in every strategy.
 *   4 PSc with the polarity correction: per step, every interior cell votes
 *     +1 / -1 / 0 by the sign of set = -sgn(vzp_s) sgn(rvxs) (vxp_s rvxs +
 *     vzp_s rvzs) (set_sign_basedon_polarization_ps, elastic_2D_kernel_2.cu:
 *     941-982; rvs = v_r - vP_r; the case vzp_s < 0, rvxs = 0 that the
 *     reference leaves undefined votes 0), the votes are counted over the 9x9
 *     window (filter_sign_new_share, :875-919), the majority sets the cell's
 *     sign (compare_sign, :921-939: +1 wins ties with either other class;
 *     a tie of -1 and 0 keeps the previous step's sign), and
 *     PScp += sign div_s curl_r.
 * div / curl are the decom stencils (elastic_2D_kernel_2.cu:621-666).
 * ------------------------------------------------------------------------- */
template <typename T>
struct XcorrRTM {
    Oracle2D<T>& o;
    explicit XcorrRTM(Oracle2D<T>& oo) : o(oo) {}
    using Model = typename Oracle2D<T>::Model;
    enum { VX, VZ, XX, ZZ, XZ, TP, VXP, VZP };

    /* Fh / Eh (optional): F / E of the step over the model region, as the
     * forward kernels store them (stress() / velocity() of Oracle2D) */
    void rvelocity(std::vector<T>* fv, const Model& m, T* Fh = nullptr) {
        const ptrdiff_t sX = o.nnz;
        const size_t mn = (size_t)o.p.nx * o.p.nz;
#pragma omp parallel for schedule(static)
        for (int i = o.R; i < o.nnx - o.R; ++i)
            for (int k = o.R; k < o.nnz - o.R; ++k) {
                size_t n = o.id(i, k);
                T px = o.cx * o.dm(fv[TP].data() + n, sX), pz = o.cz * o.dp(fv[TP].data() + n, 1);
                T f0x = o.cx * o.dm(fv[XX].data() + n, sX), f0z = o.cz * o.dm(fv[XZ].data() + n, 1);
                T f1x = o.cx * o.dp(fv[XZ].data() + n, sX), f1z = o.cz * o.dp(fv[ZZ].data() + n, 1);
                T F0 = o.pml(4, n, 0, 1, i, f0x) + o.pml(5, n, 1, 0, k, f0z);
                T F1 = o.pml(6, n, 0, 0, i, f1x) + o.pml(7, n, 1, 1, k, f1z);
                px = o.pml(8, n, 0, 1, i, px);
                pz = o.pml(9, n, 1, 1, k, pz);
                size_t q;
                if (Fh && o.model_cell(i, k, &q)) { Fh[q] = F0; Fh[mn + q] = F1; }
                T b, a = o.ab(n, &b), bo = m.B[n];
                fv[VX][n] = a * (b * fv[VX][n] + bo * -F0);
                fv[VZ][n] = a * (b * fv[VZ][n] + bo * -F1);
                fv[VXP][n] = a * (b * fv[VXP][n] + bo * -px);
                fv[VZP][n] = a * (b * fv[VZP][n] + bo * -pz);
            }
    }
    void rstress(std::vector<T>* fv, std::vector<T> (*rv)[4], const Model& m, int it, T* Eh = nullptr) {
        const ptrdiff_t sX = o.nnz;
        const size_t mn = (size_t)o.p.nx * o.p.nz;
        /* the wavelet enters where the forward step adds it, with the opposite sign */
        if (it >= 0 && it < (int)o.wavelet.size()) { fv[XX][o.src] -= o.wavelet[it]; fv[ZZ][o.src] -= o.wavelet[it]; }
#pragma omp parallel for schedule(static)
        for (int i = o.R; i < o.nnx - o.R; ++i)
            for (int k = o.R; k < o.nnz - o.R; ++k) {
                size_t n = o.id(i, k);
                T e0 = o.cx * o.dp(fv[VX].data() + n, sX), e1 = o.cz * o.dm(fv[VZ].data() + n, 1);
                T e2x = o.cx * o.dm(fv[VZ].data() + n, sX), e2z = o.cz * o.dp(fv[VX].data() + n, 1);
                T E[3] = {o.pml(0, n, 0, 0, i, e0), o.pml(2, n, 1, 0, k, e1),
                          o.pml(1, n, 0, 1, i, e2x) + o.pml(3, n, 1, 1, k, e2z)};
                size_t q;
                if (Eh && o.model_cell(i, k, &q)) { Eh[q] = E[0]; Eh[mn + q] = E[1]; Eh[2 * mn + q] = E[2]; }
                T b, a = o.ab(n, &b);
                T M = m.M[n], U = m.U[n], La = M - 2 * U;
                T sp = 0, sxx = 0, szz = 0, sxz = 0;
                for (int l = 0; l < o.L; ++l) {
                    T bl = b - m.decay[l][n], kP = m.kP[l][n], kS = m.kS[l][n];
                    T rp = a * (bl * rv[l][0][n] + kP * (E[0] + E[1]));
                    T rxx = a * (bl * rv[l][1][n] - 2 * kS * E[1]);
                    T rzz = a * (bl * rv[l][2][n] - 2 * kS * E[0]);
                    T rxz = a * (bl * rv[l][3][n] + kS * E[2]);
                    rv[l][0][n] = rp; rv[l][1][n] = rxx; rv[l][2][n] = rzz; rv[l][3][n] = rxz;
                    sp += rp; sxx += rxx; szz += rzz; sxz += rxz;
                }
                const T ne0 = -E[0], ne1 = -E[1], ne2 = -E[2], dq = -o.dtr;
                fv[XX][n] = a * (b * fv[XX][n] + M * ne0 + La * ne1 + dq * (sp + sxx));
                fv[ZZ][n] = a * (b * fv[ZZ][n] + La * ne0 + M * ne1 + dq * (sp + szz));
                fv[TP][n] = a * (b * fv[TP][n] + M * (ne0 + ne1) + dq * sp);
                fv[XZ][n] = a * (b * fv[XZ][n] + U * ne2 + dq * sxz);
            }
    }
    /* the ring: cells outside the model region shrunk by one cell, within R + 1
     * of it (R for the reversed stencils, one more for the div / curl of the
     * model region's border cells) */
    std::vector<size_t> ring() const {
        const int i0 = o.p.bl + 1, i1 = o.p.bl + o.p.nx - 1, k0 = o.p.bu + 1, k1 = o.p.bu + o.p.nz - 1;
        std::vector<size_t> out;
        for (int i = std::max(i0 - o.R - 1, 0); i < std::min(i1 + o.R + 1, o.nnx); ++i)
            for (int k = std::max(k0 - o.R - 1, 0); k < std::min(k1 + o.R + 1, o.nnz); ++k)
                if (i < i0 || i >= i1 || k < k0 || k >= k1) out.push_back(o.id(i, k));
        return out;
    }
    void div_curl(const std::vector<T>* fv, std::vector<T>& dv, std::vector<T>& cu) const {
        dv.assign(o.N, 0); cu.assign(o.N, 0);
        const ptrdiff_t sX = o.nnz;
        for (int i = o.R; i < o.nnx - o.R; ++i)
            for (int k = o.R; k < o.nnz - o.R; ++k) {
                size_t n = o.id(i, k);
                dv[n] = (T)(1.0 / o.p.dx) * o.dp(fv[VX].data() + n, sX) + (T)(1.0 / o.p.dz) * o.dm(fv[VZ].data() + n, 1);
                cu[n] = (T)(1.0 / o.p.dz) * o.dp(fv[VX].data() + n, 1) - (T)(1.0 / o.p.dx) * o.dm(fv[VZ].data() + n, sX);
            }
    }

    int run(int strategy, float* out) {
        auto& p = o.p;
        const Model& m = o.mod[1];
        if (!m.have) return -4;
        if (strategy < 0 || strategy > 2) return -1;
        if (strategy == 2 && o.L != 0) return -1;
        /* each state keeps its own C-PML memory variables: the receiver's live
         * in psiR and are swapped into the oracle around its steps */
        std::vector<T> psiR[10];
        for (auto& v : psiR) v.assign(o.cpml ? o.N : 0, (T)0);
        auto swap_psi = [&]() { for (int q = 0; q < 10; ++q) std::swap(o.psi[q], psiR[q]); };
        const size_t mn = (size_t)p.nx * p.nz;
        const int ltf = p.lt_forward < 0 ? p.lt : std::min(p.lt_forward, p.lt);
        std::vector<size_t> rg = ring();
        const size_t nr = rg.size();
        std::vector<T> hist, ringv, rings;
        if (strategy == 0) hist.assign((size_t)ltf * 3 * mn, 0);
        if (strategy == 2) { ringv.assign((size_t)ltf * nr * 4, 0); rings.assign((size_t)ltf * nr * 4, 0); }
        const int vfield[4] = {VX, VZ, VXP, VZP}, sfield[4] = {XX, ZZ, XZ, TP};
        std::vector<T> dv, cu;
        o.zero();
        for (int it = 0; it < ltf; ++it) {
            o.add_source(o.f, it);
            o.stress(o.f, o.r, m);
            if (strategy == 2)
                for (int q = 0; q < 4; ++q)
                    for (size_t j = 0; j < nr; ++j) rings[((size_t)it * 4 + q) * nr + j] = o.f[sfield[q]][rg[j]];
            o.velocity(o.f, m, 0, it);
            if (strategy == 2)
                for (int q = 0; q < 4; ++q)
                    for (size_t j = 0; j < nr; ++j) ringv[((size_t)it * 4 + q) * nr + j] = o.f[vfield[q]][rg[j]];
            if (strategy == 0) {
                div_curl(o.f, dv, cu);
                T* h = hist.data() + (size_t)it * 3 * mn;
                for (size_t q = 0; q < mn; ++q) {
                    size_t n = o.id((int)(q / p.nz) + p.bl, (int)(q % p.nz) + p.bu);
                    h[q] = o.f[VXP][n]; h[mn + q] = o.f[VZP][n]; h[2 * mn + q] = dv[n];
                }
            }
        }
        /* receiver state */
        std::vector<T> g[8];
        std::vector<T> gr[3][4];
        for (auto& v : g) v.assign(o.N, 0);
        for (int l = 0; l < o.L; ++l) for (auto& v : gr[l]) v.assign(o.N, 0);
        std::vector<double> img[5];
        for (auto& v : img) v.assign(mn, 0.0);
        std::vector<int> sign(o.N, 0), vote(o.N, 0);
        std::vector<T> prod(o.N, 0), rdv, rcu;
        const T qs = p.qcomp ? (T)-1 : (T)1;
        for (int it = p.lt - 1; it >= 0; --it) {
            for (int q = 0; q < p.nrec; ++q) {
                g[VX][o.recs[q]] += o.rec[((size_t)0 * p.lt + it) * p.nrec + q];
                g[VZ][o.recs[q]] += o.rec[((size_t)1 * p.lt + it) * p.nrec + q];
            }
            swap_psi();
            o.qsign = qs;
            o.stress(g, gr, m);
            o.qsign = 1;
            o.velocity(g, m, 0, it);
            swap_psi();
            if (it >= ltf) continue;
            /* source fields after forward step it */
            std::vector<T> sxp(o.N, 0), szp(o.N, 0), sdv(o.N, 0);
            if (strategy == 0) {
                const T* h = hist.data() + (size_t)it * 3 * mn;
                for (size_t q = 0; q < mn; ++q) {
                    size_t n = o.id((int)(q / p.nz) + p.bl, (int)(q % p.nz) + p.bu);
                    sxp[n] = h[q]; szp[n] = h[mn + q]; sdv[n] = h[2 * mn + q];
                }
            } else {
                div_curl(o.f, dv, cu);
                for (size_t q = 0; q < mn; ++q) {
                    size_t n = o.id((int)(q / p.nz) + p.bl, (int)(q % p.nz) + p.bu);
                    sxp[n] = o.f[VXP][n]; szp[n] = o.f[VZP][n]; sdv[n] = dv[n];
                }
            }
            div_curl(g, rdv, rcu);
            for (int i = o.R; i < o.nnx - o.R; ++i)
                for (int k = o.R; k < o.nnz - o.R; ++k) {
                    size_t n = o.id(i, k);
                    const T rxs = g[VX][n] - g[VXP][n], rzs = g[VZ][n] - g[VZP][n];
                    const T dot = sxp[n] * rxs + szp[n] * rzs;
                    T set = 0;
                    if (szp[n] >= 0 && rxs >= 0) set = -dot;
                    else if (szp[n] >= 0 && rxs < 0) set = dot;
                    else if (szp[n] < 0 && rxs > 0) set = dot;
                    else if (szp[n] < 0 && rxs < 0) set = -dot;
                    vote[n] = set > 0 ? 1 : (set < 0 ? 2 : 3);
                    prod[n] = sdv[n] * rcu[n];
                    size_t q;
                    if (o.model_cell(i, k, &q)) {
                        img[0][q] += sxp[n] * g[VXP][n] + szp[n] * g[VZP][n];
                        img[1][q] += dot;
                        img[2][q] += sdv[n] * rdv[n];
                        img[3][q] += prod[n];
                    }
                }
            for (int i = o.R; i < o.nnx - o.R; ++i)
                for (int k = o.R; k < o.nnz - o.R; ++k) {
                    int cnt[4] = {0, 0, 0, 0};
                    for (int di = -4; di <= 4; ++di)
                        for (int dk = -4; dk <= 4; ++dk) {
                            const int ii = i + di, kk = k + dk;
                            if (o.interior(ii, kk)) ++cnt[vote[o.id(ii, kk)]];
                        }
                    size_t n = o.id(i, k);
                    if (cnt[1] >= cnt[2] && cnt[1] >= cnt[3]) sign[n] = 1;
                    else if (cnt[2] > cnt[1] && cnt[2] > cnt[3]) sign[n] = -1;
                    else if (cnt[3] > cnt[1] && cnt[3] > cnt[2]) sign[n] = 0;
                    size_t q;
                    if (o.model_cell(i, k, &q)) img[4][q] += (T)sign[n] * prod[n];
                }
            /* undo source step it */
            if (strategy != 0) {
                if (strategy == 2)
                    for (int q = 0; q < 4; ++q)
                        for (size_t j = 0; j < nr; ++j) o.f[sfield[q]][rg[j]] = rings[((size_t)it * 4 + q) * nr + j];
                rvelocity(o.f, m);
                if (strategy == 2)
                    for (int q = 0; q < 4; ++q)
                        for (size_t j = 0; j < nr; ++j)
                            o.f[vfield[q]][rg[j]] = it > 0 ? ringv[((size_t)(it - 1) * 4 + q) * nr + j] : (T)0;
                rstress(o.f, o.r, m, it);
            }
        }
        for (int a = 0; a < 5; ++a)
            for (size_t q = 0; q < mn; ++q) out[a * mn + q] = (float)img[a][q];
        return 0;
    }
};

/* ---------------------------------------------------------------------------
 * Exact-transpose adjoint with the background from boundary strips
 * (adjoint_source = 2; no reference counterpart, elastic media only): the
 * background pass keeps only the ring of XcorrRTM::ring() after every half
 * step; stepping back from lt-1, the ring is restored before each reversed
 * half step (stress side of step it, then velocity side of step it-1, zeros
 * before step 0) and the reversed steps store F and E of step it, which
 * then drive the same transposed Born steps as the stored history.
 * ------------------------------------------------------------------------- */
template <typename T>
int adjoint_strips(Oracle2D<T>& o, const float* rr, float* gout) {
    if (o.cpml || o.L != 0) return -1;
    const auto& m = o.mod[1];
    if (!m.have) return -4;
    XcorrRTM<T> x(o);
    enum { VX, VZ, XX, ZZ, XZ, TP, VXP, VZP };
    const int vfield[4] = {VX, VZ, VXP, VZP}, sfield[4] = {XX, ZZ, XZ, TP};
    const auto& p = o.p;
    const size_t mn = (size_t)p.nx * p.nz;
    const std::vector<size_t> rg = x.ring();
    const size_t nr = rg.size();
    std::vector<T> rings((size_t)p.lt * 4 * nr), ringv((size_t)p.lt * 4 * nr);
    o.zero();
    for (int it = 0; it < p.lt; ++it) {
        o.add_source(o.f, it);
        o.stress(o.f, o.r, m);
        for (int q = 0; q < 4; ++q)
            for (size_t j = 0; j < nr; ++j) rings[((size_t)it * 4 + q) * nr + j] = o.f[sfield[q]][rg[j]];
        o.velocity(o.f, m, 0, it);
        for (int q = 0; q < 4; ++q)
            for (size_t j = 0; j < nr; ++j) ringv[((size_t)it * 4 + q) * nr + j] = o.f[vfield[q]][rg[j]];
    }
    std::vector<T> EH((size_t)p.lt * 3 * mn, 0), FH((size_t)p.lt * 2 * mn, 0);
    for (int it = p.lt - 1; it >= 0; --it) {
        for (int q = 0; q < 4; ++q)
            for (size_t j = 0; j < nr; ++j) o.f[sfield[q]][rg[j]] = rings[((size_t)it * 4 + q) * nr + j];
        x.rvelocity(o.f, m, FH.data() + (size_t)it * 2 * mn);
        for (int q = 0; q < 4; ++q)
            for (size_t j = 0; j < nr; ++j) o.f[vfield[q]][rg[j]] = it > 0 ? ringv[((size_t)(it - 1) * 4 + q) * nr + j] : (T)0;
        x.rstress(o.f, o.r, m, it, EH.data() + (size_t)it * 3 * mn);
    }
    return o.adjoint(rr, gout, &EH, &FH);
}

template <typename T>
struct Box { Oracle2D<T> o; };

struct H2 { int prec; Oracle2D<double>* d = nullptr; Oracle2D<float>* s = nullptr; };

template <typename Fn>
int disp(void* h, Fn&& fn) {
    if (!h) return -1;
    H2* H = (H2*)h;
    return H->prec == 64 ? fn(*H->d) : fn(*H->s);
}

}  // namespace

extern "C" {

void* oracle2d_create(const erm2d_params* p, int prec) {
    H2* H = new H2();
    H->prec = prec == 32 ? 32 : 64;
    int rc;
    if (H->prec == 64) { H->d = new Oracle2D<double>(); rc = H->d->init(p); }
    else { H->s = new Oracle2D<float>(); rc = H->s->init(p); }
    if (rc) { delete H->d; delete H->s; delete H; return nullptr; }
    return H;
}
void oracle2d_destroy(void* h) {
    if (!h) return;
    H2* H = (H2*)h;
    delete H->d; delete H->s; delete H;
}
int oracle2d_load_model(void* h, int which, const float* vp, const float* vs, const float* rho, const float* qp,
                        const float* qs) {
    return disp(h, [&](auto& o) { return o.load_model(which, vp, vs, rho, qp, qs); });
}
int oracle2d_set_shot(void* h, int sx, int sz) { return disp(h, [&](auto& o) { return o.set_shot(sx, sz); }); }
int oracle2d_zero(void* h, int what) {
    return disp(h, [&](auto& o) {
        if (what & 1) o.zero();
        if (what & 2) {
            std::fill(o.ex_time.begin(), o.ex_time.end(), -1);
            std::fill(o.ex_tp_time.begin(), o.ex_tp_time.end(), -1);
            for (auto* v : {&o.ex_amp, &o.ex_tp, &o.ex_vxp, &o.ex_vzp, &o.r_vxp, &o.r_vzp, &o.ex_tp_amp})
                std::fill(v->begin(), v->end(), 0);
        }
        if (what & 4) for (auto* v : {&o.img_tp, &o.img_pp, &o.img_ps}) std::fill(v->begin(), v->end(), 0);
        return 0;
    });
}
int oracle2d_step(void* h, int pass, int it_begin, int nsteps) {
    return disp(h, [&](auto& o) { return o.step(pass, it_begin, nsteps); });
}
int oracle2d_get_field(void* h, int fid, double* dst) {
    return disp(h, [&](auto& o) {
        using T = typename std::remove_reference<decltype(o.c[0])>::type;
        if (fid >= 0 && fid < 8) { for (size_t n = 0; n < o.N; ++n) dst[n] = o.f[fid][n]; return 0; }
        if (fid == ERM2D_DIV || fid == ERM2D_CURL) {
            std::vector<T> dv, cu;
            o.decom(dv, cu);
            const auto& s = fid == ERM2D_DIV ? dv : cu;
            for (size_t n = 0; n < o.N; ++n) dst[n] = s[n];
            return 0;
        }
        if (fid == ERM2D_PDIV || fid == ERM2D_SCURL) {
            std::vector<T> dv, cu;
            o.decom_new(o.last_model, dv, cu);
            const auto& s = fid == ERM2D_PDIV ? dv : cu;
            for (size_t n = 0; n < o.N; ++n) dst[n] = s[n];
            return 0;
        }
        if (fid == ERM2D_EX_TIME) { for (size_t n = 0; n < o.N; ++n) dst[n] = o.ex_time[n]; return 0; }
        if (fid == ERM2D_EX_TP_TIME) { for (size_t n = 0; n < o.N; ++n) dst[n] = o.ex_tp_time[n]; return 0; }
        const std::vector<T>* s = fid == ERM2D_EX_TP_AMP ? &o.ex_tp_amp : fid == ERM2D_EX_AMP ? &o.ex_amp : fid == ERM2D_IMG_TP ? &o.img_tp
                                : fid == ERM2D_IMG_PP ? &o.img_pp : fid == ERM2D_IMG_PS ? &o.img_ps
                                : fid == ERM2D_EX_TP ? &o.ex_tp : fid == ERM2D_EX_VXP ? &o.ex_vxp
                                : fid == ERM2D_EX_VZP ? &o.ex_vzp : fid == ERM2D_R_VXP ? &o.r_vxp
                                : fid == ERM2D_R_VZP ? &o.r_vzp : nullptr;
        if (!s) return -1;
        for (size_t n = 0; n < o.N; ++n) dst[n] = (*s)[n];
        return 0;
    });
}
int oracle2d_set_field(void* h, int fid, const double* src) {
    return disp(h, [&](auto& o) {
        if (fid < 0 || fid >= 8) return -1;
        for (size_t n = 0; n < o.N; ++n) o.f[fid][n] = src[n];
        return 0;
    });
}
int oracle2d_record(void* h, double* dst, const double* src) {
    return disp(h, [&](auto& o) {
        if (src) for (size_t n = 0; n < o.rec.size(); ++n) o.rec[n] = src[n];
        if (dst) for (size_t n = 0; n < o.rec.size(); ++n) dst[n] = o.rec[n];
        return 0;
    });
}
/* the imaging condition of step it on the current fields, no propagation */
int oracle2d_image(void* h, int it) {
    return disp(h, [&](auto& o) { o.image(o.f, it); return 0; });
}
int oracle2d_born(void* h, const float* dm, float* d) { return disp(h, [&](auto& o) { return o.born(dm, d); }); }
int oracle2d_adjoint(void* h, const float* r, float* g) { return disp(h, [&](auto& o) { return o.adjoint(r, g); }); }

int oracle2d_adjoint_strips(void* h, const float* r, float* g) {
    return disp(h, [&](auto& o) { return adjoint_strips(o, r, g); });
}
/* cross-correlation RTM (records as set by oracle2d_record); out [5][nx][nz] */
int oracle2d_rtm_xcorr(void* h, int strategy, float* out) {
    return disp(h, [&](auto& o) {
        using T = typename std::remove_reference<decltype(o.c[0])>::type;
        XcorrRTM<T> X(o);
        return X.run(strategy, out);
    });
}
/* reference-mode adjoint: g = (g_lambda, g_mu, g_rho) [3][nx][nz] */
int oracle2d_adjoint_ref(void* h, const float* r, float* g) {
    return disp(h, [&](auto& o) { return adjoint_ref(o, r, g); });
}
/* reference-mode adjoint propagation alone: start from av = (vx, vz) with zero
 * adjoint stresses and memory, run nsteps reversed steps without records, and
 * return av (out [2][nnx*nnz]) and the moduli-weighted stresses of the
 * adjoint state S = (Mp as_xx + (Mp-2Ms) as_zz + sum_l P_l ar_xx + (P_l-2S_l) ar_zz,
 * the zz counterpart, Ms as_xz + sum_l S_l ar_xz) (out [3][nnx*nnz]) */
int oracle2d_ref_adjoint_steps(void* h, const double* vx, const double* vz, int nsteps, double* vout, double* sout) {
    return disp(h, [&](auto& o) {
        const auto& m = o.mod[1];
        if (!m.have) return -4;
        using T = typename std::remove_reference<decltype(o.c[0])>::type;
        RefAdjoint<T> A(o);
        for (size_t n = 0; n < o.N; ++n) { A.av[0][n] = (T)vx[n]; A.av[1][n] = (T)vz[n]; }
        for (int s = 0; s < nsteps; ++s) A.step(m);
        for (size_t n = 0; n < o.N; ++n) {
            vout[n] = A.av[0][n]; vout[o.N + n] = A.av[1][n];
            const T Bn = m.B[n], Mp = m.M[n] * Bn, Ms = m.U[n] * Bn;
            T sx = Mp * A.as[0][n] + (Mp - 2 * Ms) * A.as[1][n], sz = (Mp - 2 * Ms) * A.as[0][n] + Mp * A.as[1][n];
            T sxz = Ms * A.as[2][n];
            for (int l = 0; l < o.L; ++l) {
                const T P = m.kP[l][n] * Bn, S = m.kS[l][n] * Bn;
                sx += P * A.ar[l][0][n] + (P - 2 * S) * A.ar[l][1][n];
                sz += (P - 2 * S) * A.ar[l][0][n] + P * A.ar[l][1][n];
                sxz += S * A.ar[l][2][n];
            }
            sout[n] = sx; sout[o.N + n] = sz; sout[2 * o.N + n] = sxz;
        }
        return 0;
    });
}
/* SLS parameter check: returns fM, fU, gP[0..2], gS[0..2], taus[0..2] for one cell */
int oracle2d_sls(int L, double freq, double qp, double qs, double* out) {
    Sls s = sls_params(L, freq, qp, qs);
    out[0] = s.fM; out[1] = s.fU;
    for (int l = 0; l < 3; ++l) { out[2 + l] = s.gP[l]; out[5 + l] = s.gS[l]; out[8 + l] = s.taus[l]; }
    return 0;
}

}  // extern "C"
